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

// sink velocity of condensate jc (1-based) in layer jl; radii from mo_dictionary.F90:475-527
__device__ __forceinline__ double v_fall(const Ctx &c, const StateDev &old, int jc, int ji, int jk, int jl) {
  switch (jc) {
    case 1:
      return v_fall_solid(500.e-6);
    case 2:
    case 5:  // radius of rain drops and graupel: diag%a_rain, which the reference keeps at zero
      return v_fall_liquid(c, old, 0.0, ji, jk, jl);
    case 3:
      return v_fall_solid(95.e-6);
    case 4:
      return v_fall_liquid(c, old, 12.e-6, ji, jk, jl);
    default:
      return 0.0;
  }
}

// three_band_solver_gen_densities, mo_column_solvers.F90:353-604; all tracers of a column in one
// lane so the sediment heating accumulates over the constituents in the reference's order
__global__ void __launch_bounds__(BCOL) k_solve_gen(const Ctx c, const StateDev old, const StateDev nw, int rk) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, ncond = c.G.ncond;
  if (col >= nx * ny) return;
  const int jk = col % nx + 1, ji = col / nx + 1;
  const double itw = 0.5, ew = 1.0 - itw;
  const double dt = c.C.dtime;
  double *sep = c.w.col[0], *srp = c.w.col[1];
  for (int jc = 1; jc <= c.G.nc; ++jc) {
    if (jc == ncond + 1) continue;
    const double *ro = old.rho[jc - 1];
    const double *tr = c.w.tend_rho[jc - 1];
    // interface fluxes of interface q (between layers q and q+1)
    auto iface = [&](int q, double &fi, double &fr, double &fe) {
      const size_t iw = LG_IX(c.G, ji, jk, q + 1);
      fi = old.wind_w[iw];
      fr = nw.wind_w[iw];
      double vf = 0.0;
      if (jc <= 5) vf = 0.5 * (v_fall(c, old, jc, ji, jk, q) + v_fall(c, old, jc, ji, jk, q + 1));
      fi = fi - vf;
      fr = fr - vf;
      fi = c.g.area_z[iw] * fi;
      fr = c.g.area_z[iw] * fr;
      double dens;
      const double tmp = AT(c.w.temperature, ji, jk, q + 1);
      if (fr >= 0.0)
        dens = AT(ro, ji, jk, q + 1);
      else
        dens = AT(ro, ji, jk, q);
      fr = dens * fr;
      fe = c_p_cond(jc, tmp) * tmp * fr;
    };
    double fi_up = 0.0, fr_up = 0.0, fe_up = 0.0;  // interface above the current layer
    double fi_dn, fr_dn, fe_dn;                    // interface below
    double ep_prev = 0.0, rp_prev = 0.0, c_prev = 0.0;
    for (int jl = 1; jl <= nl; ++jl) {
      const size_t i = LG_IX(c.G, ji, jk, jl);
      const double V = c.g.volume[i];
      if (jl < nl) iface(jl, fi_dn, fr_dn, fe_dn);
      if (rk == 1 && jc == 1) c.w.csh[i] = 0.0;
      double d, e = 0.0, cq = 0.0;
      if (jl == 1) {
        d = (fi_dn >= 0.0) ? 1.0 : 1.0 - itw * dt / V * fi_dn;
      } else if (jl == nl) {
        d = (fi_up >= 0.0) ? 1.0 + itw * dt / V * fi_up : 1.0;
        if (jc <= ncond) d = d + itw * v_fall(c, old, jc, ji, jk, nl) * dt * AT(c.g.area_z, ji, jk, nl + 1) / V;
      } else {
        d = 1.0;
        if (fi_up >= 0.0) d = d + itw * dt / V * fi_up;
        if (fi_dn < 0.0) d = d - itw * dt / V * fi_dn;
      }
      if (jl < nl) {  // off-diagonals of row jl (mo_column_solvers.F90:487-495)
        if (fi_dn >= 0.0) {
          cq = 0.0;
          e = -itw * dt / V * fi_dn;
        } else {
          cq = itw * dt / c.g.volume[LG_IX(c.G, ji, jk, jl + 1)] * fi_dn;
          e = 0.0;
        }
      }
      double r = ro[i] + dt * tr[i];
      if (jl == 1) {
        r = r + ew * dt * fr_dn / V;
        if (rk == 1 && jc <= ncond) c.w.csh[i] = c.w.csh[i] + fe_dn / V;
      } else if (jl == nl) {
        r = r - ew * dt * fr_up / V;
        if (rk == 1 && jc <= ncond) c.w.csh[i] = c.w.csh[i] - fe_up / V;
        if (jc <= ncond) {
          const double vfs = v_fall(c, old, jc, ji, jk, nl);
          r = r - ew * vfs * dt * ro[i] * AT(c.g.area_z, ji, jk, nl + 1) / V;
          if (jc > ncond / 4 && rk == 1) {
            const double tn = AT(c.w.temperature, ji, jk, nl);
            c.w.csh[i] = c.w.csh[i] - vfs * tn * c_p_cond(jc, tn) * ro[i] * AT(c.g.area_z, ji, jk, nl + 1) / V;
          }
        }
      } else {
        r = r + ew * dt * (-fr_up + fr_dn) / V;
        if (rk == 1 && jc <= ncond) c.w.csh[i] = c.w.csh[i] + (-fe_up + fe_dn) / V;
      }
      double epq, rpq;
      if (jl == 1) {
        epq = e / d;
        rpq = r / d;
      } else {
        const double den = d - ep_prev * c_prev;
        epq = e / den;
        rpq = (r - rp_prev * c_prev) / den;
      }
      sep[i] = epq;
      srp[i] = rpq;
      ep_prev = epq;
      rp_prev = rpq;
      c_prev = cq;
      fi_up = fi_dn;
      fr_up = fr_dn;
      fe_up = fe_dn;
    }
    double x = srp[LG_IX(c.G, ji, jk, nl)];
    double *rn = nw.rho[jc - 1];
    rn[LG_IX(c.G, ji, jk, nl)] = x < 0.0 ? 0.0 : x;
    for (int jl = nl - 1; jl >= 1; --jl) {
      const size_t i = LG_IX(c.G, ji, jk, jl);
      x = srp[i] - sep[i] * x;
      rn[i] = x < 0.0 ? 0.0 : x;
    }
  }
}

}  // namespace lgame
