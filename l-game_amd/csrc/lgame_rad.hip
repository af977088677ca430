// lgame_rad.hip -- radiation boundary of the HEVI step: the coupler steps of
// calc_radiative_flux_convergence (src/radiation/mo_rterrtmgp_coupler.F90:56-493) with the gray
// two-stream solver of include/lgame_gpu.h in the place of RTE+RRTMGP, the surrogate's column
// features, and the application and audit of the surrogate's outputs.
//
// One column per lane (consecutive lanes take consecutive x, so the per-level loads of a
// wavefront are coalesced); per-level intermediates live in the handle's FP64 column scratch,
// which is free at the start of a step.  All radiation arithmetic is FP64 in both forms; the
// result goes to radiation_tendency in the stencil type R.
#include "lgame_ops.h"
#include "lgame_rad.h"
#include "lgame_step.h"

#include "../../include/lgame_gpu.h"

namespace lgame {

// coszenith, mo_rterrtmgp_coupler.F90:550-629
__device__ double coszenith(double lat, double lon, double t) {
  const double omega = 7.292115e-5, omega_rev = 1.99099e-7, obliquity = 0.409092592843564;
  const double t_0 = 1576843200.0, phi_0_earth_around_sun = 0.0, phi_0_earth_rotation = 0.0;
  const double t_transformed = t - t_0;
  const double rot_angle = omega * t_transformed - phi_0_earth_rotation;
  const double ne0 = cos(lat) * cos(lon), ne1 = cos(lat) * sin(lon), ne2 = sin(lat);
  const double cr = cos(rot_angle), sr = sin(rot_angle), co = cos(obliquity), so = sin(obliquity);
  // trans_earth2sun applied to the normal vector (rows of the matrix times ne)
  const double ns0 = (-cr * co) * ne0 + (sr * co) * ne1 + so * ne2;
  const double ns1 = (-sr) * ne0 + (-cr) * ne1 + 0.0 * ne2;
  const double ns2 = (cr * so) * ne0 + (-sr * so) * ne1 + co * ne2;
  const double s0 = cos(omega_rev * t_transformed + phi_0_earth_around_sun);
  const double s1 = sin(omega_rev * t_transformed + phi_0_earth_around_sun);
  double cz = ns0 * (-s0) + ns1 * (-s1) + ns2 * (-0.0);
  if (cz < 0.0) cz = 0.0;
  return cz;
}

__device__ __forceinline__ double rad_clip(double t) {
  if (t > LGAME_RAD_T_MAX) t = LGAME_RAD_T_MAX;
  if (t < LGAME_RAD_T_MIN) t = LGAME_RAD_T_MIN;
  return t;
}

__device__ __forceinline__ unsigned rad_hash(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

template <class R>
__global__ void __launch_bounds__(BCOL) k_rad(const CtxT<R> c, const StateDev s, const RadDev r, double t_rad, int flags) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev, ncond = c.G.ncond;
  if (col >= nx * ny) return;
  const int jk = col % nx + 1, ji = col / nx + 1;
  const bool audit = flags & RAD_AUDIT;
  const int nin = 5 * nl + 4, nout = nl + 2;
  // a prediction outside the training envelope (or not finite) always goes back to the solver
  bool outside = false;
  if (audit && r.y_scale) {
    const float *yc = r.ypred + (size_t)col * nout;
    for (int o = 0; o < nout; ++o)
      if (!(fabsf(yc[o] - r.y_mean[o]) <= (float)LGAME_RAD_ENVELOPE * r.y_scale[o])) outside = true;
  }
  const bool sampled = rad_hash((unsigned)col ^ (unsigned)(r.call * 0x9E3779B9ull)) < r.audit_thresh;
  if (audit && !outside && !sampled) return;
  // per-level scratch: temperature, pressure, vapour/liquid/ice mixing ratios, water paths,
  // interface pressure, LW transmissivity, source, SW optical depth, SW heating
  const Fld<double> sT = c.w.col[0], sp = c.w.col[1], sqv = c.w.col[2], slwp = c.w.col[3], siwp = c.w.col[4];
  const Fld<double> spi = c.w.col[5], stl = c.w.col[6], sB = c.w.col[7], sts = c.w.col[8], sq = c.w.col[9];
  const Fld<double> sdn = c.w.col[10], sup = c.w.col[11];
  const Fld<const double> rg = s.rho[ncond];
  float *x = (flags & RAD_X) ? r.x + (size_t)col * nin : nullptr;
  // mo_rterrtmgp_coupler.F90:158-240: equation-of-state pressure and the water paths
  for (int jl = 1; jl <= nl; ++jl) {
    const ix_t i = LG_IX(c.G, ji, jk, jl);
    const double T = (double)c.w.temperature[i];
    const double rho_g = rg[i];
    const double p = R_D * rho_g * T;
    const double thickness = AT(c.g64.z_w, ji, jk, jl) - AT(c.g64.z_w, ji, jk, jl + 1);
    double qv = 0.0, ql = 0.0, qi = 0.0, lwp = 0.0, iwp = 0.0;
    if (c.C.lmoist) {
      const double r1 = s.rho[0][i], r2 = s.rho[1][i], r3 = s.rho[2][i], r4 = s.rho[3][i], r5 = s.rho[4][i];
      iwp = thickness * 1000.0 * (r1 + r3 + r5);
      lwp = thickness * 1000.0 * (r2 + r4);
      qv = s.rho[ncond + 1][i] / rho_g;
      qi = (r1 + r3 + r5) / rho_g;
      ql = (r2 + r4) / rho_g;
    }
    const double Tc = rad_clip(T);
    sT[i] = Tc;
    sp[i] = p;
    sqv[i] = qv;
    slwp[i] = lwp;
    siwp[i] = iwp;
    if (x) {
      x[jl - 1] = (float)Tc;
      x[nl + jl - 1] = (float)p;
      x[2 * nl + jl - 1] = (float)qv;
      x[3 * nl + jl - 1] = (float)ql;
      x[4 * nl + jl - 1] = (float)qi;
    }
  }
  // :259-296 interface pressures (the TOA value uses z_scalar of layer 2, as the reference)
  for (int jl = 1; jl <= nlev; ++jl) {
    double pv;
    if (jl == 1)
      pv = sp[LG_IX(c.G, ji, jk, 1)] * exp(-(AT(c.g64.z_w, ji, jk, 1) - AT(c.g64.z_scalar, ji, jk, 2)) / LGAME_RAD_SCALE_HEIGHT);
    else if (jl == nlev)
      pv = sp[LG_IX(c.G, ji, jk, nl)] * exp(-(AT(c.g64.z_w, ji, jk, nlev) - AT(c.g64.z_scalar, ji, jk, nl)) / LGAME_RAD_SCALE_HEIGHT);
    else
      pv = 0.5 * (sp[LG_IX(c.G, ji, jk, jl - 1)] + sp[LG_IX(c.G, ji, jk, jl)]);
    spi[LG_IX(c.G, ji, jk, jl)] = pv;
  }
  const double mu0 = coszenith(r.lat[col], r.lon[col], t_rad);
  const double albedo = r.albedo[col];
  double t_sfc_in = r.tsfc[col];
  if (c.C.lsoil) {  // temp_sfc_full of the sea and soil surface, mo_manage_radiation_calls.F90:36-42
    const ix_t i1 = LG_IX(c.G, ji, jk, 1), dl = (ix_t)c.G.lvl;
    const double sea = 1.0 - c.g64.land_fraction[i1] - c.g64.soil[i1];
    t_sfc_in = sea * c.g64.soil[i1 + 4 * dl] + (1.0 - sea) * s.temperature_soil[i1];
  }
  const double t_sfc = rad_clip(t_sfc_in);
  const double p_sfc = spi[LG_IX(c.G, ji, jk, nlev)];
  if (x) {
    x[5 * nl] = (float)t_sfc;
    x[5 * nl + 1] = (float)mu0;
    x[5 * nl + 2] = (float)albedo;
    x[5 * nl + 3] = (float)p_sfc;
  }
  if (!(flags & (RAD_TEND | RAD_Y | RAD_AUDIT))) return;
  // optical depths, transmissivities and sources of the gray two-stream solver
  for (int jl = 1; jl <= nl; ++jl) {
    const ix_t i = LG_IX(c.G, ji, jk, jl);
    const double dm = (spi[LG_IX(c.G, ji, jk, jl + 1)] - spi[i]) / LGAME_RAD_GRAVITY;
    const double qv = sqv[i], lwp = slwp[i], iwp = siwp[i];
    const double tau = dm * (LGAME_RAD_K_LW_DRY * sp[i] / LGAME_RAD_P_0 + LGAME_RAD_K_LW_VAPOUR * qv) + LGAME_RAD_K_LW_LIQUID * lwp +
                       LGAME_RAD_K_LW_ICE * iwp;
    const double tau_sw =
        dm * (LGAME_RAD_K_SW_DRY + LGAME_RAD_K_SW_VAPOUR * qv) + LGAME_RAD_K_SW_LIQUID * lwp + LGAME_RAD_K_SW_ICE * iwp;
    stl[i] = exp(-LGAME_RAD_DIFFUSIVITY * tau);
    const double T = sT[i], t2 = T * T;
    sB[i] = LGAME_RAD_SIGMA * (t2 * t2);
    sts[i] = tau_sw;
    sq[i] = 0.0;
  }
  // shortwave on day columns; calc_power_density (:495-547) adds SW, then LW
  double sw = 0.0;
  if (mu0 > 0.0) {
    double dn = LGAME_RAD_SOLAR_CONSTANT * mu0;
    sdn[LG_IX(c.G, ji, jk, 1)] = dn;
    for (int jl = 1; jl <= nl; ++jl) {
      dn = dn * exp(-sts[LG_IX(c.G, ji, jk, jl)] / mu0);
      sdn[LG_IX(c.G, ji, jk, jl + 1)] = dn;
    }
    double up = albedo * dn;
    sw = dn - up;
    sup[LG_IX(c.G, ji, jk, nlev)] = up;
    for (int jl = nl; jl >= 1; --jl) {
      up = up * exp(-LGAME_RAD_DIFFUSIVITY * sts[LG_IX(c.G, ji, jk, jl)]);
      sup[LG_IX(c.G, ji, jk, jl)] = up;
    }
    for (int jl = 1; jl <= nl; ++jl) {
      const ix_t i = LG_IX(c.G, ji, jk, jl), ib = LG_IX(c.G, ji, jk, jl + 1);
      sq[i] = sq[i] + (sup[ib] - sup[i] + sdn[i] - sdn[ib]) / (AT(c.g64.z_w, ji, jk, jl) - AT(c.g64.z_w, ji, jk, jl + 1));
    }
  }
  {
    double dn = 0.0;
    sdn[LG_IX(c.G, ji, jk, 1)] = dn;
    for (int jl = 1; jl <= nl; ++jl) {
      const ix_t i = LG_IX(c.G, ji, jk, jl);
      dn = dn * stl[i] + sB[i] * (1.0 - stl[i]);
      sdn[LG_IX(c.G, ji, jk, jl + 1)] = dn;
    }
    const double ts2 = t_sfc * t_sfc;
    double up = LGAME_RAD_EMISSIVITY * LGAME_RAD_SIGMA * (ts2 * ts2) + (1.0 - LGAME_RAD_EMISSIVITY) * dn;
    sup[LG_IX(c.G, ji, jk, nlev)] = up;
    const double lw = up - dn;
    for (int jl = nl; jl >= 1; --jl) {
      const ix_t i = LG_IX(c.G, ji, jk, jl);
      up = up * stl[i] + sB[i] * (1.0 - stl[i]);
      sup[i] = up;
    }
    double se = 0.0;
    const float *yp = audit ? r.ypred + (size_t)col * nout : nullptr;
    float *y = (flags & RAD_Y) ? r.y + (size_t)col * nout : nullptr;
    for (int jl = 1; jl <= nl; ++jl) {
      const ix_t i = LG_IX(c.G, ji, jk, jl), ib = LG_IX(c.G, ji, jk, jl + 1);
      const double q = sq[i] + (sup[ib] - sup[i] + sdn[i] - sdn[ib]) / (AT(c.g64.z_w, ji, jk, jl) - AT(c.g64.z_w, ji, jk, jl + 1));
      sq[i] = q;
      const double hr = 86400.0 * q / (rg[i] * C_D_P);
      if (y) y[jl - 1] = (float)hr;
      if (yp) {
        const double e = (double)yp[jl - 1] - hr;
        se = se + e * e;
      }
    }
    if (y) {
      y[nl] = (float)sw;
      y[nl + 1] = (float)lw;
    }
    bool write = (flags & RAD_TEND) != 0;
    if (audit) {
      const double rmse = sqrt(se / nl);
      if (sampled && !outside) {  // the error statistic of the random sample inside the envelope
        atomicAdd(r.counts + 0, 1ull);
        atomicAdd(r.sq, se);
      }
      if (outside || rmse > r.tol) {  // this column goes back to the solver
        atomicAdd(r.counts + 1, 1ull);
        write = true;
      }
    }
    if (write) {
      for (int jl = 1; jl <= nl; ++jl) {
        const ix_t i = LG_IX(c.G, ji, jk, jl);
        c.w.radiation_tendency[i] = (R)sq[i];
      }
      r.sw_in[col] = sw;
      r.lw_out[col] = lw;
    }
  }
}

// radiation_tendency and surface fluxes from the surrogate's outputs (heating rates in K/day)
template <class R>
__global__ void __launch_bounds__(BCOL) k_rad_apply(const CtxT<R> c, const StateDev s, const RadDev r) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (col >= nx * ny) return;
  const int jk = col % nx + 1, ji = col / nx + 1;
  const float *yp = r.ypred + (size_t)col * (nl + 2);
  const Fld<const double> rg = s.rho[c.G.ncond];
  for (int jl = 1; jl <= nl; ++jl) {
    const ix_t i = LG_IX(c.G, ji, jk, jl);
    c.w.radiation_tendency[i] = (R)((double)yp[jl - 1] * rg[i] * C_D_P / 86400.0);
  }
  r.sw_in[col] = yp[nl];
  r.lw_out[col] = yp[nl + 1];
}

template <class R>
void enqueue_radiation(const CtxT<R> &c, const StateDev &old, const RadCall &rc, hipStream_t st) {
  const dim3 g((c.G.nx * c.G.ny + BCOL - 1) / BCOL), b(BCOL);
  if (rc.mode == LGAME_RAD_TEACHER) {
    hipLaunchKernelGGL(k_rad<R>, g, b, 0, st, c, old, rc.dev, rc.t_rad, RAD_TEND);
  } else {
    hipLaunchKernelGGL(k_rad<R>, g, b, 0, st, c, old, rc.dev, rc.t_rad, RAD_X);
    mlp_infer_device(rc.mlp, rc.dev.x, (long long)c.G.nx * c.G.ny, const_cast<float *>(rc.dev.ypred), st);
    hipLaunchKernelGGL(k_rad_apply<R>, g, b, 0, st, c, old, rc.dev);
    hipLaunchKernelGGL(k_rad<R>, g, b, 0, st, c, old, rc.dev, rc.t_rad, RAD_AUDIT);
  }
}

template void enqueue_radiation<double>(const CtxT<double> &, const StateDev &, const RadCall &, hipStream_t);
template void enqueue_radiation<float>(const CtxT<float> &, const StateDev &, const RadCall &, hipStream_t);
template __global__ void k_rad<double>(const CtxT<double>, const StateDev, const RadDev, double, int);
template __global__ void k_rad<float>(const CtxT<float>, const StateDev, const RadDev, double, int);

}  // namespace lgame
