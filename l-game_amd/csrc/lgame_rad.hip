// lgame_rad.hip -- radiation boundary of the HEVI step: the coupler steps of
// calc_radiative_flux_convergence (src/radiation/mo_rterrtmgp_coupler.F90:56-493) with the gray
// two-stream solver of include/lgame_gpu.h in the place of RTE+RRTMGP, the surrogate's column
// features, and the application and audit of the surrogate's outputs.
//
// One column per lane (consecutive lanes take consecutive x, so the per-level loads of a
// wavefront are coalesced); per-level intermediates live in the handle's FP64 column scratch,
// which is free at the start of a step.  All radiation arithmetic is FP64 in both forms; the
// result goes to radiation_tendency in the stencil type R.
//
// The surrogate path: k_rad_gather forms the network's column inputs on its LGAME_RAD_LEVELS
// levels (a column of another layer count is resampled linearly in the layer index) and writes
// them row-major through an LDS tile, the network runs on the handle's own buffers, k_rad_apply
// maps its heating rates back to the column's layers, k_rad_select lists the audit sample and the
// columns outside the training envelope, and k_rad re-runs the solver on that list only.
#include "lgame_ops.h"
#include "lgame_rad.h"
#include "lgame_step.h"
#include "lgame_batch.h"

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

__device__ __forceinline__ bool rad_sampled(const RadDev &r, int col) {
  return rad_hash((unsigned)col ^ (unsigned)(r.call * 0x9E3779B9ull)) < r.audit_thresh;
}

// The network sees LGAME_RAD_LEVELS levels.  Level q of the network lies at layer position
// q (nl - 1)/(L - 1) of a column of nl layers (0-based), layer l at network position
// l (L - 1)/(nl - 1): the index below and the weight of the one above, linear in the index.
__device__ __forceinline__ void rad_position(int i, int from, int to, int &lo, double &w) {
  if (from == to) {
    lo = i;
    w = 0.0;
    return;
  }
  const double t = (double)i * (double)(to - 1) / (double)(from - 1);
  lo = (int)t;
  if (lo > to - 2) lo = to - 2;
  w = t - (double)lo;
}

// the surrogate's heating rate [K/day] at layer jl (1-based) of a column of nl layers
__device__ __forceinline__ double rad_pred_layer(const float *yp, int jl, int nl) {
  constexpr int L = LGAME_RAD_LEVELS;
  if (nl == L) return (double)yp[jl - 1];
  int q;
  double w;
  rad_position(jl - 1, nl, L, q, w);
  return (1.0 - w) * (double)yp[q] + w * (double)yp[q + 1];
}

template <class R>
__device__ __forceinline__ void b_rad(const CtxT<R> &c, const StateDev &s, const RadDev &r, double t_rad, int flags, const Bid bid) {
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev, ncond = c.G.ncond;
  const bool audit = flags & RAD_AUDIT;
  int col = bid.x * blockDim.x + threadIdx.x;
  bool outside = false, sampled = false;
  if (audit) {  // the columns k_rad_select listed: the audit sample and those outside the envelope
    if (col >= (int)*r.list_count) return;
    const int e = r.list[col];
    outside = e < 0;
    col = outside ? ~e : e;
    sampled = rad_sampled(r, col);
  } else if (col >= nx * ny) {
    return;
  }
  const int jk = col % nx + 1, ji = col / nx + 1;
  const int nin = 5 * nl + 4, nout = nl + 2;
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
    const float *yp = audit ? r.ypred + (size_t)col * LGAME_RAD_N_OUT : nullptr;
    float *y = (flags & RAD_Y) ? r.y + (size_t)col * nout : nullptr;
    for (int jl = 1; jl <= nl; ++jl) {
      const ix_t i = LG_IX(c.G, ji, jk, jl), ib = LG_IX(c.G, ji, jk, jl + 1);
      const double q = sq[i] + (sup[ib] - sup[i] + sdn[i] - sdn[ib]) / (AT(c.g64.z_w, ji, jk, jl) - AT(c.g64.z_w, ji, jk, jl + 1));
      sq[i] = q;
      const double hr = 86400.0 * q / (rg[i] * C_D_P);
      if (y) y[jl - 1] = (float)hr;
      if (yp) {
        const double e = rad_pred_layer(yp, jl, nl) - hr;
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
      // the reference stores the surface shortwave flux of day columns only, night columns keep
      // the previous call's value (mo_rterrtmgp_coupler.F90:430-432)
      if (mu0 > 0.0) r.sw_in[col] = sw;
      r.lw_out[col] = lw;
    }
  }
}
template <class R>
__global__ void __launch_bounds__(BCOL) k_rad(const CtxT<R> c, const StateDev s, const RadDev r, double t_rad, int flags) {
  b_rad(c, s, r, t_rad, flags, bid_here());
}

// the surrogate's column inputs (T, p, q_v, q_l, q_i on the network's levels, then the surface
// temperature, the cosine of the zenith angle, the albedo and the surface pressure), the same
// quantities k_rad forms with RAD_X.  Each lane forms its column's row in an LDS tile; the block
// then writes its rows, which are contiguous in x, as one coalesced range.
constexpr int RAD_TILE_LD = LGAME_RAD_N_IN + 1;  // odd row stride: the lanes' rows hit different banks

template <class R>
__device__ __forceinline__ void b_rad_gather(const CtxT<R> &c, const StateDev &s, const RadDev &r, double t_rad, const Bid bid) {
  constexpr int L = LGAME_RAD_LEVELS;
  __shared__ float tile[BCOL * RAD_TILE_LD];
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev, ncond = c.G.ncond;
  const int ncol = nx * ny, col0 = bid.x * BCOL, col = col0 + (int)threadIdx.x;
  if (bid.x == 0 && threadIdx.x == 0) *r.list_count = 0u;  // k_rad_select appends after this launch
  if (col < ncol) {
    const int jk = col % nx + 1, ji = col / nx + 1;
    float *row = tile + threadIdx.x * RAD_TILE_LD;
    const Fld<const double> rg = s.rho[ncond];
    // the five per-layer inputs of layer jl, as k_rad forms them
    auto feat = [&](int jl, double f[5]) {
      const ix_t i = LG_IX(c.G, ji, jk, jl);
      const double T = (double)c.w.temperature[i];
      const double rho_g = rg[i];
      f[0] = rad_clip(T);
      f[1] = R_D * rho_g * T;
      f[2] = f[3] = f[4] = 0.0;
      if (c.C.lmoist) {
        const double r1 = s.rho[0][i], r2 = s.rho[1][i], r3 = s.rho[2][i], r4 = s.rho[3][i], r5 = s.rho[4][i];
        f[2] = s.rho[ncond + 1][i] / rho_g;
        f[3] = (r2 + r4) / rho_g;
        f[4] = (r1 + r3 + r5) / rho_g;
      }
    };
    for (int q = 0; q < L; ++q) {
      double f[5];
      if (nl == L) {
        feat(q + 1, f);
      } else {
        int lo;
        double w, g[5];
        rad_position(q, L, nl, lo, w);
        feat(lo + 1, f);
        feat(lo + 2, g);
#pragma unroll
        for (int k = 0; k < 5; ++k) f[k] = (1.0 - w) * f[k] + w * g[k];
      }
#pragma unroll
      for (int k = 0; k < 5; ++k) row[k * L + q] = (float)f[k];
    }
    double t_sfc_in = r.tsfc[col];
    if (c.C.lsoil) {  // temp_sfc_full of the sea and soil surface, mo_manage_radiation_calls.F90:36-42
      const ix_t i1 = LG_IX(c.G, ji, jk, 1), dl = (ix_t)c.G.lvl;
      const double sea = 1.0 - c.g64.land_fraction[i1] - c.g64.soil[i1];
      t_sfc_in = sea * c.g64.soil[i1 + 4 * dl] + (1.0 - sea) * s.temperature_soil[i1];
    }
    const ix_t ib = LG_IX(c.G, ji, jk, nl);
    const double p_low = R_D * rg[ib] * (double)c.w.temperature[ib];
    const double p_sfc = p_low * exp(-(AT(c.g64.z_w, ji, jk, nlev) - AT(c.g64.z_scalar, ji, jk, nl)) / LGAME_RAD_SCALE_HEIGHT);
    row[5 * L] = (float)rad_clip(t_sfc_in);
    row[5 * L + 1] = (float)coszenith(r.lat[col], r.lon[col], t_rad);
    row[5 * L + 2] = (float)r.albedo[col];
    row[5 * L + 3] = (float)p_sfc;
  }
  __syncthreads();
  const int nrows = ncol - col0 < BCOL ? ncol - col0 : BCOL;
  float *dst = r.x + (size_t)col0 * LGAME_RAD_N_IN;
  for (int e = threadIdx.x; e < nrows * LGAME_RAD_N_IN; e += BCOL) {
    const int rw = e / LGAME_RAD_N_IN, k = e - rw * LGAME_RAD_N_IN;
    dst[e] = tile[rw * RAD_TILE_LD + k];
  }
}
template <class R>
__global__ void __launch_bounds__(BCOL) k_rad_gather(const CtxT<R> c, const StateDev s, const RadDev r, double t_rad) {
  b_rad_gather(c, s, r, t_rad, bid_here());
}

// radiation_tendency and surface fluxes from the surrogate's outputs (heating rates in K/day on
// the network's levels, mapped to the column's layers)
template <class R>
__device__ __forceinline__ void b_rad_apply(const CtxT<R> &c, const StateDev &s, const RadDev &r, double t_rad, const Bid bid) {
  constexpr int L = LGAME_RAD_LEVELS;
  const int col = bid.x * blockDim.x + threadIdx.x;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (col >= nx * ny) return;
  const int jk = col % nx + 1, ji = col / nx + 1;
  const float *yp = r.ypred + (size_t)col * LGAME_RAD_N_OUT;
  const Fld<const double> rg = s.rho[c.G.ncond];
  for (int jl = 1; jl <= nl; ++jl) {
    const ix_t i = LG_IX(c.G, ji, jk, jl);
    c.w.radiation_tendency[i] = (R)(rad_pred_layer(yp, jl, nl) * rg[i] * C_D_P / 86400.0);
  }
  // day columns only, as the solver path (mo_rterrtmgp_coupler.F90:430-432)
  if (coszenith(r.lat[col], r.lon[col], t_rad) > 0.0) r.sw_in[col] = yp[L];
  r.lw_out[col] = yp[L + 1];
}
template <class R>
__global__ void __launch_bounds__(BCOL) k_rad_apply(const CtxT<R> c, const StateDev s, const RadDev r, double t_rad) {
  b_rad_apply(c, s, r, t_rad, bid_here());
}

// the audit list: every column of the hashed sample and every column whose prediction leaves
// the training envelope (or is not finite), appended in wavefront batches
__device__ __forceinline__ void b_rad_select(const RadDev &r, int ncol, const Bid bid) {
  const int col = bid.x * blockDim.x + threadIdx.x;
  bool take = false, outside = false;
  if (col < ncol) {
    const float *yc = r.ypred + (size_t)col * LGAME_RAD_N_OUT;
    for (int o = 0; o < LGAME_RAD_N_OUT; ++o)
      if (!(fabsf(yc[o] - r.y_mean[o]) <= (float)LGAME_RAD_ENVELOPE * r.y_scale[o])) outside = true;
    take = outside || rad_sampled(r, col);
  }
  const unsigned long long mask = __ballot(take);  // every lane of the wavefront takes part
  if (!mask) return;
  const int leader = __ffsll((long long)mask) - 1;
  const int lane = (int)__lane_id();
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(r.list_count, (unsigned)__popcll(mask));
  base = __shfl(base, leader);
  if (take) {
    const unsigned below = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
    r.list[base + below] = outside ? ~col : col;
  }
}
__global__ void __launch_bounds__(BCOL) k_rad_select(const RadDev r, int ncol) { b_rad_select(r, ncol, bid_here()); }

template <class R>
int enqueue_radiation(const CtxT<R> &c, const StateDev &old, const RadCall &rc, hipStream_t st) {
  const int ncol = c.G.nx * c.G.ny;
  const dim3 g((ncol + BCOL - 1) / BCOL), b(BCOL);
  if (rc.mode == LGAME_RAD_TEACHER) {
    hipLaunchKernelGGL(k_rad<R>, g, b, 0, st, c, old, rc.dev, rc.t_rad, RAD_TEND);
  } else {
    hipLaunchKernelGGL(k_rad_gather<R>, g, b, 0, st, c, old, rc.dev, rc.t_rad);
    const int rc_inf = mlp_infer_device(rc.mlp, rc.work, rc.dev.x, (long long)ncol, const_cast<float *>(rc.dev.ypred), st);
    if (rc_inf) return rc_inf;
    hipLaunchKernelGGL(k_rad_apply<R>, g, b, 0, st, c, old, rc.dev, rc.t_rad);
    hipLaunchKernelGGL(k_rad_select, g, b, 0, st, rc.dev, ncol);
    hipLaunchKernelGGL(k_rad<R>, g, b, 0, st, c, old, rc.dev, rc.t_rad, RAD_AUDIT);
  }
  return hipGetLastError() == hipSuccess ? 0 : LGAME_EHIP;
}

template int enqueue_radiation<double>(const CtxT<double> &, const StateDev &, const RadCall &, hipStream_t);
template int enqueue_radiation<float>(const CtxT<float> &, const StateDev &, const RadCall &, hipStream_t);
template __global__ void k_rad<double>(const CtxT<double>, const StateDev, const RadDev, double, int);
template __global__ void k_rad<float>(const CtxT<float>, const StateDev, const RadDev, double, int);

// ---- the same kernels over a group of domains (one launch per phase) ------------------------
LG_BATCH_KERNEL(rad_solver, BCOL, PH_RAD, b_rad(d.c, d.s[a.old_slot], d.rad, d.t_rad, a.flags, bid))
LG_BATCH_KERNEL(rad_gather, BCOL, PH_RAD, b_rad_gather(d.c, d.s[a.old_slot], d.rad, d.t_rad, bid))
LG_BATCH_KERNEL(rad_apply, BCOL, PH_RAD, b_rad_apply(d.c, d.s[a.old_slot], d.rad, d.t_rad, bid))
LG_BATCH_KERNEL(rad_select, BCOL, PH_RAD, b_rad_select(d.rad, d.c.G.nx * d.c.G.ny, bid))

template <class R>
void batch_launch_radiation(int rph, dim3 g, dim3 b, hipStream_t st, const MemberDev<R> *M, const int *prefix, int nm,
                            const BatchArgs &a) {
  switch (rph) {
    case RPH_SOLVER: hipLaunchKernelGGL(kb_rad_solver<R>, g, b, 0, st, M, prefix, nm, a); break;
    case RPH_GATHER: hipLaunchKernelGGL(kb_rad_gather<R>, g, b, 0, st, M, prefix, nm, a); break;
    case RPH_APPLY: hipLaunchKernelGGL(kb_rad_apply<R>, g, b, 0, st, M, prefix, nm, a); break;
    case RPH_SELECT: hipLaunchKernelGGL(kb_rad_select<R>, g, b, 0, st, M, prefix, nm, a); break;
    default: break;
  }
}
template void batch_launch_radiation<double>(int, dim3, dim3, hipStream_t, const MemberDev<double> *, const int *, int, const BatchArgs &);
template void batch_launch_radiation<float>(int, dim3, dim3, hipStream_t, const MemberDev<float> *, const int *, int, const BatchArgs &);

}  // namespace lgame
