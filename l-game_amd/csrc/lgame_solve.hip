// lgame_solve.hip -- implicit vertical column solvers of the HEVI step (FP64, gfx950).
//
// One column per lane: consecutive lanes take consecutive x, so every per-level load and store
// of a wavefront is one coalesced 512-byte row segment.  The tridiagonal coefficients of a
// column are built while streaming down the column with a rolling window of three layers; the
// forward Thomas elimination runs in the same sweep, and only the eliminated coefficients plus
// the few per-level values the back substitution needs go to coalesced column scratch.
#include "lgame_ops.h"
#include "lgame_step.h"

namespace lgame {

// per-layer quantities of three_band_solver_ver, mo_column_solvers.F90:136-174
struct Layer {
  double rho_o, rt_o, ep_o, eb, z, V, al, be, ga, rhoe, rte, tpe, epe, ratio;
};

__device__ __forceinline__ Layer load_layer(const Ctx &c, const StateDev &old, const StateDev &nu, int rk, size_t i) {
  const int g = c.G.ncond;
  const double dt = c.C.dtime;
  const double pw = 0.5;
  Layer L;
  L.rho_o = old.rho[g][i];
  L.rt_o = old.rhotheta_v[i];
  L.ep_o = old.exner_pert[i];
  L.eb = c.g.exner_bg[i];
  L.z = c.g.z_scalar[i];
  L.V = c.g.volume[i];
  const double trho = c.w.tend_rho[g][i], trt = c.w.tend_rhotheta_v[i];
  L.rhoe = L.rho_o + dt * trho;
  L.rte = L.rt_o + dt * trt;
  if (rk == 1) {
    L.al = -L.rt_o / (L.rho_o * L.rho_o) / L.V;
    L.be = 1.0 / L.rho_o / L.V;
    L.ga = R_D / (C_D_V * L.rt_o) * (L.eb + L.ep_o) / L.V;
    L.ratio = L.rt_o / L.rho_o;
  } else {
    const double rn = nu.rho[g][i], rtn = nu.rhotheta_v[i], epn = nu.exner_pert[i];
    const double a_old = -L.rt_o / (L.rho_o * L.rho_o);
    const double b_old = 1.0 / L.rho_o;
    const double g_old = R_D / (C_D_V * L.rt_o) * (L.eb + L.ep_o);
    const double a_new = -rtn / (rn * rn);
    const double b_new = 1.0 / rn;
    const double g_new = R_D / (C_D_V * rtn) * (L.eb + epn);
    L.al = ((1.0 - pw) * a_old + pw * a_new) / L.V;
    L.be = ((1.0 - pw) * b_old + pw * b_new) / L.V;
    L.ga = ((1.0 - pw) * g_old + pw * g_new) / L.V;
    L.ratio = rtn / rn;
  }
  L.tpe = old.theta_v_pert[i] + dt * L.V * (L.al * trho + L.be * trt);
  L.epe = L.ep_o + dt * L.V * L.ga * trt;
  return L;
}

// interface quantities between layers q and q+1 (mo_column_solvers.F90:177-184, 190-196)
struct Iface {
  double rio, rie, tin, dp, az, wo, tw;
};

__device__ __forceinline__ Iface make_iface(const Ctx &c, const StateDev &old, const Layer &a, const Layer &b, size_t iw) {
  Iface f;
  f.rio = 0.5 * (a.rho_o + b.rho_o);
  f.rie = 0.5 * (a.rhoe + b.rhoe);
  f.tin = 0.5 * (a.ratio + b.ratio);
  f.dp = c.g.klemp_pref[iw];
  f.az = c.g.area_z[iw];
  f.wo = old.wind_w[iw];
  f.tw = c.w.tend_w[iw];
  return f;
}

// three_band_solver_ver, mo_column_solvers.F90:21-351 (soil and surface heat flux off)
__global__ void __launch_bounds__(BCOL) k_solve_ver(const Ctx c, const StateDev old, const StateDev nw, int rk) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, g = c.G.ncond;
  if (col >= nx * ny) return;
  const int jk = col % nx + 1, ji = col / nx + 1;
  const StateDev &nu = rk == 1 ? old : nw;
  const double dt = c.C.dtime;
  double *sep = c.w.col[0], *srp = c.w.col[1], *srhoe = c.w.col[2], *srte = c.w.col[3];
  double *stin = c.w.col[4], *sga = c.w.col[5], *srio = c.w.col[6];
  const double K2 = IMPL_THERMO_WEIGHT * (dt * dt);
  const double K1 = IMPL_THERMO_WEIGHT * dt * C_D_P;

  // forward sweep: rows q = 1..nl-1 of the interface mass-flux system
  Layer L0 = load_layer(c, old, nu, rk, LG_IX(c.G, ji, jk, 1));
  Layer L1 = load_layer(c, old, nu, rk, LG_IX(c.G, ji, jk, 2));
  Iface F0 = make_iface(c, old, L0, L1, LG_IX(c.G, ji, jk, 2));
  Layer L2 = L1;
  Iface F1 = F0;
  double ep_prev = 0.0, rp_prev = 0.0, c_prev = 0.0;
  {
    const size_t i0 = LG_IX(c.G, ji, jk, 1);
    srhoe[i0] = L0.rhoe;
    srte[i0] = L0.rte;
    sga[i0] = L0.ga;
  }
  for (int q = 1; q <= nl - 1; ++q) {
    // layer q+1 = L1 (stored for the back substitution); layer q+2 and interface q+1 if present
    {
      const size_t i1 = LG_IX(c.G, ji, jk, q + 1);
      srhoe[i1] = L1.rhoe;
      srte[i1] = L1.rte;
      sga[i1] = L1.ga;
    }
    const bool has_next = q + 2 <= nl;
    if (has_next) {
      L2 = load_layer(c, old, nu, rk, LG_IX(c.G, ji, jk, q + 2));
      F1 = make_iface(c, old, L1, L2, LG_IX(c.G, ji, jk, q + 2));
    }
    const double dzs = L0.z - L1.z;
    const double deb = L0.eb - L1.eb;
    const double d = -F0.tin * F0.tin * (L0.ga + L1.ga) + 0.5 * deb * (L1.al - L0.al + F0.tin * (L1.be - L0.be)) -
                     dzs / (K2 * C_D_P * F0.rio) * (2.0 / F0.az + dt * F0.wo * 0.5 * (-1.0 / L0.V + 1.0 / L1.V)) * F0.dp;
    const double r = -(F0.wo + dt * F0.tw) * dzs / (K2 * C_D_P) + F0.tin * (L0.epe - L1.epe) / dt + 0.5 / dt * (L0.tpe + L1.tpe) * deb -
                     dzs / (K2 * C_D_P) * F0.wo * F0.rie / F0.rio;
    double cq = 0.0, eq = 0.0;
    if (has_next) {
      cq = F1.tin * L1.ga * F0.tin + 0.5 * (L1.eb - L2.eb) * (L1.al + L1.be * F0.tin) -
           (L1.z - L2.z) / K1 * 0.5 * F1.wo / (L1.V * F1.rio) * F1.dp;
      eq = F0.tin * L1.ga * F1.tin - 0.5 * deb * (L1.al + L1.be * F1.tin) + dzs / K1 * 0.5 * F0.wo / (L1.V * F0.rio) * F0.dp;
    }
    // thomas_algorithm, mo_column_solvers.F90:622-632
    double epq, rpq;
    if (q == 1) {
      epq = eq / d;
      rpq = r / d;
    } else {
      const double den = d - ep_prev * c_prev;
      epq = eq / den;
      rpq = (r - rp_prev * c_prev) / den;
    }
    const size_t iq = LG_IX(c.G, ji, jk, q);
    sep[iq] = epq;
    srp[iq] = rpq;
    stin[iq] = F0.tin;
    srio[iq] = F0.rio;
    ep_prev = epq;
    rp_prev = rpq;
    c_prev = cq;
    L0 = L1;
    L1 = L2;
    F0 = F1;
  }

  // back substitution (mo_column_solvers.F90:635-638) and the new state (:305-349)
  const int n = nl - 1;
  double x_low = srp[LG_IX(c.G, ji, jk, n)];  // x(n)
  double rho_below = 0.0;
  for (int Lr = nl; Lr >= 1; --Lr) {
    const size_t i = LG_IX(c.G, ji, jk, Lr);
    const double V = c.g.volume[i];
    double rho_new, rt_new;
    double x_up = 0.0;
    if (Lr == nl) {
      x_up = x_low;  // x(nl-1)
      rho_new = srhoe[i] - dt * x_up / V;
      rt_new = srte[i] - dt * stin[LG_IX(c.G, ji, jk, nl - 1)] * x_up / V;
    } else if (Lr >= 2) {
      const size_t iu = LG_IX(c.G, ji, jk, Lr - 1);
      x_up = srp[iu] - sep[iu] * x_low;
      rho_new = srhoe[i] + dt * (-x_up + x_low) / V;
      rt_new = srte[i] + dt * (-stin[iu] * x_up + stin[i] * x_low) / V;
    } else {
      rho_new = srhoe[i] + dt * x_low / V;
      rt_new = srte[i] + dt * stin[i] * x_low / V;
    }
    nw.rho[g][i] = rho_new;
    nw.rhotheta_v[i] = rt_new;
    if (Lr < nl) {  // w at the interface below this layer, level Lr+1
      const size_t ib = LG_IX(c.G, ji, jk, Lr + 1);
      const double rin = 0.5 * (rho_new + rho_below);
      nw.wind_w[ib] = (2.0 * x_low / c.g.area_z[ib] - rin * old.wind_w[ib]) / srio[i];
    }
    nw.exner_pert[i] = old.exner_pert[i] + V * sga[i] * (rt_new - old.rhotheta_v[i]);
    nw.theta_v_pert[i] = rt_new / rho_new - c.g.theta_v_bg[i];
    rho_below = rho_new;
    x_low = x_up;
  }
}

// v_fall_liquid / v_fall_solid, mo_derived.F90:218-283; calc_diff_coeff :189-216
__device__ double v_fall_liquid(const Ctx &c, const StateDev &s, double radius, int ji, int jk, int jl) {
  const double reynolds_crit = 10.0;
  double gl;
  if (jl == 1)
    gl = AT(c.g.gravity_m_v, ji, jk, jl + 1);
  else if (jl == c.G.nl)
    gl = AT(c.g.gravity_m_v, ji, jk, jl);
  else
    gl = AT(c.g.ipw[4], ji, jk, jl) * AT(c.g.gravity_m_v, ji, jk, jl) + AT(c.g.ipw[5], ji, jk, jl) * AT(c.g.gravity_m_v, ji, jk, jl + 1);
  const double rho = AT(s.rho[c.G.ncond], ji, jk, jl);
  const double particle_radius = 130e-12;
  const double particle_mass = M_D / N_A;
  const double temperature = AT(c.w.temperature, ji, jk, jl);
  const double thermal_velocity = sqrt(8.0 * K_B * temperature / (MPI * particle_mass));
  const double particle_density = rho / particle_mass;
  const double cross_section = 4.0 * MPI * (particle_radius * particle_radius);
  const double mean_free_path = 1.0 / (sqrt(2.0) * particle_density * cross_section);
  const double kin = (1.0 / 3.0) * thermal_velocity * mean_free_path;
  double v = 2.0 * (radius * radius) * (RHO_H2O - rho) * gl / (9.0 * rho * kin);
  const double re = v * radius / kin;
  if (re > reynolds_crit) {
    const double c_w = 0.8;
    v = sqrt(8.0 * radius * RHO_H2O * gl / (3.0 * rho * c_w));
  }
  return v;
}

__device__ __forceinline__ double v_fall_solid(double radius) { return (radius > (double)250.e-6f) ? 1.0 : 0.001; }

// three_band_solver_gen_densities, mo_column_solvers.F90:353-604.  All tracers of a column are
// eliminated in one sweep down the column (they share the winds, metrics, temperature and fall
// speeds of each level); per level the tracers are visited in the reference's constituent order,
// so the sediment heating accumulates in the same order.  Sink velocities: snow v_fall_solid of
// 500e-6 m = 1 m/s, cloud ice v_fall_solid of 95e-6 m = 0.001 m/s, cloud water v_fall_liquid of
// 12e-6 m; rain and graupel use the radius diag%a_rain, which the reference keeps at zero, so their
// v_fall_liquid is exactly zero (mo_dictionary.F90:475-527, control.F90:443).
constexpr int NTR = 6;  // tracers: snow, rain, cloud ice, cloud water, graupel, vapour

__global__ void __launch_bounds__(BCOL) k_solve_gen(const Ctx c, const StateDev old, const StateDev nw, int rk) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, ncond = c.G.ncond;
  if (col >= nx * ny) return;
  const int jk = col % nx + 1, ji = col / nx + 1;
  const double itw = 0.5, ew = 1.0 - itw;
  const double dt = c.C.dtime;
  const int jcs[NTR] = {1, 2, 3, 4, 5, 7};  // 1-based constituent index of each tracer
  const double vf_snow = v_fall_solid(500.e-6), vf_ice = v_fall_solid(95.e-6);
  double ep[NTR], rp[NTR], cprev[NTR], fi_up[NTR], fr_up[NTR], fe_up[NTR];
#pragma unroll
  for (int t = 0; t < NTR; ++t) ep[t] = rp[t] = cprev[t] = fi_up[t] = fr_up[t] = fe_up[t] = 0.0;
  double vcw_lo = v_fall_liquid(c, old, 12.e-6, ji, jk, 1);  // cloud-water fall speed of the layer
  for (int jl = 1; jl <= nl; ++jl) {
    const size_t i = LG_IX(c.G, ji, jk, jl);
    const double V = c.g.volume[i];
    // interface below this layer (q = jl), shared by all tracers
    double wo = 0.0, wn = 0.0, az = 0.0, tif = 0.0, cpi = 0.0, cpw = 0.0, vcw_up = vcw_lo, Vb = 0.0;
    const bool below = jl < nl;
    if (below) {
      const size_t iw = LG_IX(c.G, ji, jk, jl + 1);
      wo = old.wind_w[iw];
      wn = nw.wind_w[iw];
      az = c.g.area_z[iw];
      tif = c.w.temperature[iw];  // temperature of layer jl+1 (used for both flux directions)
      cpi = c_p_ice(tif);
      cpw = c_p_water(tif);
      vcw_lo = v_fall_liquid(c, old, 12.e-6, ji, jk, jl + 1);
      Vb = c.g.volume[iw];
    }
    const double vcw_layer = vcw_up;  // v_fall of cloud water in layer jl
    double csh = 0.0;
#pragma unroll
    for (int t = 0; t < NTR; ++t) {
      const int jc = jcs[t];
      const double *ro = old.rho[jc - 1];
      double fi = 0.0, fr = 0.0, fe = 0.0;
      if (below) {
        double vf = 0.0;
        if (jc == 1)
          vf = 0.5 * (vf_snow + vf_snow);
        else if (jc == 3)
          vf = 0.5 * (vf_ice + vf_ice);
        else if (jc == 4)
          vf = 0.5 * (vcw_layer + vcw_lo);
        else if (jc == 2 || jc == 5)
          vf = 0.5 * (0.0 + 0.0);
        fi = wo - vf;
        fr = wn - vf;
        fi = az * fi;
        fr = az * fr;
        const double dens = fr >= 0.0 ? ro[LG_IX(c.G, ji, jk, jl + 1)] : ro[i];
        fr = dens * fr;
        const double cp = ((jc - 1) % 2 == 0) ? cpi : cpw;
        fe = cp * tif * fr;
      }
      double d, e = 0.0, cq = 0.0;
      if (jl == 1) {
        d = (fi >= 0.0) ? 1.0 : 1.0 - itw * dt / V * fi;
      } else if (jl == nl) {
        d = (fi_up[t] >= 0.0) ? 1.0 + itw * dt / V * fi_up[t] : 1.0;
      } else {
        d = 1.0;
        if (fi_up[t] >= 0.0) d = d + itw * dt / V * fi_up[t];
        if (fi < 0.0) d = d - itw * dt / V * fi;
      }
      double vfs = 0.0;  // sink velocity in the lowest layer
      if (jl == nl && jc <= ncond) {
        vfs = jc == 1 ? vf_snow : jc == 3 ? vf_ice : jc == 4 ? vcw_layer : 0.0;
        d = d + itw * vfs * dt * AT(c.g.area_z, ji, jk, nl + 1) / V;
      }
      if (below) {
        if (fi >= 0.0) {
          cq = 0.0;
          e = -itw * dt / V * fi;
        } else {
          cq = itw * dt / Vb * fi;
          e = 0.0;
        }
      }
      double r = ro[i] + dt * c.w.tend_rho[jc - 1][i];
      if (jl == 1) {
        r = r + ew * dt * fr / V;
        if (rk == 1 && jc <= ncond) csh = csh + fe / V;
      } else if (jl == nl) {
        r = r - ew * dt * fr_up[t] / V;
        if (rk == 1 && jc <= ncond) csh = csh - fe_up[t] / V;
        if (jc <= ncond) {
          r = r - ew * vfs * dt * ro[i] * AT(c.g.area_z, ji, jk, nl + 1) / V;
          if (jc > ncond / 4 && rk == 1) {
            const double tn = AT(c.w.temperature, ji, jk, nl);
            csh = csh - vfs * tn * c_p_cond(jc, tn) * ro[i] * AT(c.g.area_z, ji, jk, nl + 1) / V;
          }
        }
      } else {
        r = r + ew * dt * (-fr_up[t] + fr) / V;
        if (rk == 1 && jc <= ncond) csh = csh + (-fe_up[t] + fe) / V;
      }
      double epq, rpq;
      if (jl == 1) {
        epq = e / d;
        rpq = r / d;
      } else {
        const double den = d - ep[t] * cprev[t];
        epq = e / den;
        rpq = (r - rp[t] * cprev[t]) / den;
      }
      c.w.col[2 * t][i] = epq;
      c.w.col[2 * t + 1][i] = rpq;
      ep[t] = epq;
      rp[t] = rpq;
      cprev[t] = cq;
      fi_up[t] = fi;
      fr_up[t] = fr;
      fe_up[t] = fe;
    }
    if (rk == 1) c.w.csh[i] = csh;
  }
  // back substitution and clipping of round-off negatives (mo_column_solvers.F90:584-596)
  double x[NTR];
#pragma unroll
  for (int t = 0; t < NTR; ++t) x[t] = 0.0;
  for (int jl = nl; jl >= 1; --jl) {
    const size_t i = LG_IX(c.G, ji, jk, jl);
#pragma unroll
    for (int t = 0; t < NTR; ++t) {
      x[t] = (jl == nl) ? c.w.col[2 * t + 1][i] : c.w.col[2 * t + 1][i] - c.w.col[2 * t][i] * x[t];
      nw.rho[jcs[t] - 1][i] = x[t] < 0.0 ? 0.0 : x[t];
    }
  }
}

}  // namespace lgame
