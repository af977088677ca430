// lgame_solve.hip -- implicit vertical column solvers of the HEVI step (FP64 in both forms, gfx950).
//
// One column per lane: consecutive lanes take consecutive x, so every per-level load and store
// of a wavefront is one coalesced 512-byte row segment.  The tridiagonal coefficients of a
// column are built while streaming down the column with a rolling window of three layers; the
// forward Thomas elimination runs in the same sweep, and only the eliminated coefficients plus
// the few per-level values the back substitution needs go to coalesced column scratch.
#include "lgame_ops.h"
#include "lgame_step.h"

namespace lgame {

// Raw inputs of one layer k of three_band_solver_ver and of the interface at its top (level k).
// A column sweep loads them one or two layers ahead of their use, so that the loads of the
// next layers are in flight while the current row is eliminated and its scratch is stored.
struct RawLayer {
  double rho_o, rt_o, ep_o, tp_o, eb, z, V, trho, trt, rn, rtn, epn;  // layer k
  double dp, az, wo, tw;                                               // interface k
};

template <class R>
__device__ __forceinline__ RawLayer load_raw(const CtxT<R> &c, const StateDev &old, const StateDev &nu, int rk, ix_t i) {
  const int g = c.G.ncond;
  RawLayer r;
  r.rho_o = old.rho[g][i];
  r.rt_o = old.rhotheta_v[i];
  r.ep_o = old.exner_pert[i];
  r.tp_o = old.theta_v_pert[i];
  r.eb = c.g64.exner_bg[i];
  r.z = c.g64.z_scalar[i];
  r.V = c.g64.volume[i];
  r.trho = (double)c.w.tend_rho[g][i];
  r.trt = (double)c.w.tend_rhotheta_v[i];
  if (rk == 2) {
    r.rn = nu.rho[g][i];
    r.rtn = nu.rhotheta_v[i];
    r.epn = nu.exner_pert[i];
  } else {
    r.rn = r.rtn = r.epn = 0.0;
  }
  r.dp = c.g64.klemp_pref[i];
  r.az = c.g64.area_z[i];
  r.wo = old.wind_w[i];
  r.tw = (double)c.w.tend_w[i];
  return r;
}

// per-layer quantities of three_band_solver_ver, mo_column_solvers.F90:136-174
struct Layer {
  double rho_o, rt_o, ep_o, eb, z, V, al, be, ga, rhoe, rte, tpe, epe, ratio;
};

__device__ __forceinline__ Layer make_layer(const RawLayer &r, int rk, double dt) {
  const double pw = 0.5;
  Layer L;
  L.rho_o = r.rho_o;
  L.rt_o = r.rt_o;
  L.ep_o = r.ep_o;
  L.eb = r.eb;
  L.z = r.z;
  L.V = r.V;
  const double trho = r.trho, trt = r.trt;
  L.rhoe = L.rho_o + dt * trho;
  L.rte = L.rt_o + dt * trt;
  if (rk == 1) {
    L.al = -L.rt_o / (L.rho_o * L.rho_o) / L.V;
    L.be = 1.0 / L.rho_o / L.V;
    L.ga = R_D / (C_D_V * L.rt_o) * (L.eb + L.ep_o) / L.V;
    L.ratio = L.rt_o / L.rho_o;
  } else {
    const double rn = r.rn, rtn = r.rtn, epn = r.epn;
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
  L.tpe = r.tp_o + dt * L.V * (L.al * trho + L.be * trt);
  L.epe = L.ep_o + dt * L.V * L.ga * trt;
  return L;
}

// interface quantities between layers q and q+1 (mo_column_solvers.F90:177-184, 190-196);
// rb holds the interface q+1 values
struct Iface {
  double rio, rie, tin, dp, az, wo, tw;
};

__device__ __forceinline__ Iface make_iface(const Layer &a, const Layer &b, const RawLayer &rb) {
  Iface f;
  f.rio = 0.5 * (a.rho_o + b.rho_o);
  f.rie = 0.5 * (a.rhoe + b.rhoe);
  f.tin = 0.5 * (a.ratio + b.ratio);
  f.dp = rb.dp;
  f.az = rb.az;
  f.wo = rb.wo;
  f.tw = rb.tw;
  return f;
}

// inputs of one layer of the back substitution, loaded one layer ahead
struct BackIn {
  double V, rhoe, rte, ga, rio, ep_o, rt_o, tb;  // layer Lr
  double rp_u, ep_u, tin_u;                      // row Lr-1 of the system (interface Lr)
  double az_b, wo_b;                             // interface Lr+1 (below the layer)
};

// three_band_solver_ver, mo_column_solvers.F90:21-351 (soil and surface heat flux off)
template <class R>
__global__ void __launch_bounds__(BCOL) k_solve_ver(const CtxT<R> c, const StateDev old, const StateDev nw, int rk) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, g = c.G.ncond;
  if (col >= nx * ny) return;
  const int jk = col % nx + 1, ji = col / nx + 1;
  const StateDev &nu = rk == 1 ? old : nw;
  const double dt = c.C.dtime;
  const Fld<double> sep = c.w.col[0], srp = c.w.col[1], srhoe = c.w.col[2], srte = c.w.col[3];
  const Fld<double> stin = c.w.col[4], sga = c.w.col[5], srio = c.w.col[6];
  const double K2 = IMPL_THERMO_WEIGHT * (dt * dt);
  const double K1 = IMPL_THERMO_WEIGHT * dt * C_D_P;

  // forward sweep: rows q = 1..nl-1 of the interface mass-flux system
  Layer L0 = make_layer(load_raw(c, old, nu, rk, LG_IX(c.G, ji, jk, 1)), rk, dt);
  const RawLayer raw1 = load_raw(c, old, nu, rk, LG_IX(c.G, ji, jk, 2));
  RawLayer nxt = raw1;  // layer q+2 for the first row
  if (nl >= 3) nxt = load_raw(c, old, nu, rk, LG_IX(c.G, ji, jk, 3));
  Layer L1 = make_layer(raw1, rk, dt);
  Iface F0 = make_iface(L0, L1, raw1);
  Layer L2 = L1;
  Iface F1 = F0;
  double ep_prev = 0.0, rp_prev = 0.0, c_prev = 0.0;
  {
    const ix_t i0 = LG_IX(c.G, ji, jk, 1);
    srhoe[i0] = L0.rhoe;
    srte[i0] = L0.rte;
    sga[i0] = L0.ga;
  }
  for (int q = 1; q <= nl - 1; ++q) {
    const bool has_next = q + 2 <= nl;
    const RawLayer cur = nxt;  // layer q+2
    if (q + 3 <= nl) nxt = load_raw(c, old, nu, rk, LG_IX(c.G, ji, jk, q + 3));
    // layer q+1 = L1 (stored for the back substitution); layer q+2 and interface q+1 if present
    {
      const ix_t i1 = LG_IX(c.G, ji, jk, q + 1);
      srhoe[i1] = L1.rhoe;
      srte[i1] = L1.rte;
      sga[i1] = L1.ga;
    }
    if (has_next) {
      L2 = make_layer(cur, rk, dt);
      F1 = make_iface(L1, L2, cur);
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
    const ix_t iq = LG_IX(c.G, ji, jk, q);
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
  auto load_back = [&](int Lr) {
    BackIn b;
    const ix_t i = LG_IX(c.G, ji, jk, Lr);
    b.V = c.g64.volume[i];
    b.rhoe = srhoe[i];
    b.rte = srte[i];
    b.ga = sga[i];
    b.rio = Lr < nl ? srio[i] : 0.0;
    b.ep_o = old.exner_pert[i];
    b.rt_o = old.rhotheta_v[i];
    b.tb = c.g64.theta_v_bg[i];
    b.rp_u = b.ep_u = b.tin_u = 0.0;
    if (Lr >= 2) {
      const ix_t iu = LG_IX(c.G, ji, jk, Lr - 1);
      b.rp_u = srp[iu];
      b.ep_u = sep[iu];
      b.tin_u = stin[iu];
    }
    b.az_b = b.wo_b = 0.0;
    if (Lr < nl) {
      const ix_t ib = LG_IX(c.G, ji, jk, Lr + 1);
      b.az_b = c.g64.area_z[ib];
      b.wo_b = old.wind_w[ib];
    }
    return b;
  };
  double x_low = 0.0;  // x(Lr) of the row below the current layer
  double tin_low = 0.0;
  double rho_below = 0.0;
  BackIn bn = load_back(nl);
  for (int Lr = nl; Lr >= 1; --Lr) {
    const BackIn b = bn;
    if (Lr >= 2) bn = load_back(Lr - 1);
    const double az_b = b.az_b, wo_b = b.wo_b;
    const ix_t i = LG_IX(c.G, ji, jk, Lr);
    const double V = b.V;
    double rho_new, rt_new;
    double x_up = 0.0;
    if (Lr == nl) {
      x_up = b.rp_u;  // x(nl-1)
      rho_new = b.rhoe - dt * x_up / V;
      rt_new = b.rte - dt * b.tin_u * x_up / V;
    } else if (Lr >= 2) {
      x_up = b.rp_u - b.ep_u * x_low;
      rho_new = b.rhoe + dt * (-x_up + x_low) / V;
      rt_new = b.rte + dt * (-b.tin_u * x_up + tin_low * x_low) / V;
    } else {
      rho_new = b.rhoe + dt * x_low / V;
      rt_new = b.rte + dt * tin_low * x_low / V;
    }
    nw.rho[g][i] = rho_new;
    nw.rhotheta_v[i] = rt_new;
    if (Lr < nl) {  // w at the interface below this layer, level Lr+1
      const double rin = 0.5 * (rho_new + rho_below);
      nw.wind_w[LG_IX(c.G, ji, jk, Lr + 1)] = (2.0 * x_low / az_b - rin * wo_b) / b.rio;
    }
    nw.exner_pert[i] = b.ep_o + V * b.ga * (rt_new - b.rt_o);
    nw.theta_v_pert[i] = rt_new / rho_new - b.tb;
    rho_below = rho_new;
    x_low = x_up;
    tin_low = b.tin_u;
  }
}

// v_fall_liquid / v_fall_solid, mo_derived.F90:218-283; calc_diff_coeff :189-216.
// The loads (VfIn, vf_load) are separate from the arithmetic (vf_liquid) so that a column sweep
// can issue them one level ahead.
struct VfIn {
  double g0, g1, w4, w5, rho, temp;
};

template <class R>
__device__ __forceinline__ VfIn vf_load(const CtxT<R> &c, const StateDev &s, int ji, int jk, int jl) {
  VfIn v;
  v.g1 = v.w4 = v.w5 = 0.0;
  if (jl == 1) {
    v.g0 = AT(c.g64.gravity_m_v, ji, jk, jl + 1);
  } else if (jl == c.G.nl) {
    v.g0 = AT(c.g64.gravity_m_v, ji, jk, jl);
  } else {
    v.w4 = AT(c.g64.ipw[4], ji, jk, jl);
    v.g0 = AT(c.g64.gravity_m_v, ji, jk, jl);
    v.w5 = AT(c.g64.ipw[5], ji, jk, jl);
    v.g1 = AT(c.g64.gravity_m_v, ji, jk, jl + 1);
  }
  v.rho = AT(s.rho[c.G.ncond], ji, jk, jl);
  v.temp = (double)AT(c.w.temperature, ji, jk, jl);
  return v;
}

__device__ __forceinline__ double vf_liquid(const VfIn &in, double radius, int jl, int nl) {
  const double reynolds_crit = 10.0;
  const double gl = (jl == 1 || jl == nl) ? in.g0 : in.w4 * in.g0 + in.w5 * in.g1;
  const double rho = in.rho;
  const double particle_radius = 130e-12;
  const double particle_mass = M_D / N_A;
  const double temperature = in.temp;
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

// three_band_solver_gen_densities, mo_column_solvers.F90:353-604.
//
// A block holds 64 columns and one wavefront per tracer (snow, rain, cloud ice, cloud water,
// graupel, vapour), so every lane eliminates one tracer of one column and the per-level loads of
// the winds, metrics and temperature that all tracers of a column share are served from the
// CU's cache.  At the predictor substep the five condensates also contribute to the sediment
// heating, which the reference accumulates per level in constituent order; each condensate wave
// leaves its per-level terms in LDS, and after every GEN_K levels the vapour wave (which has the
// least work) adds them up in that order and stores the heating.
// Sink velocities: snow v_fall_solid of 500e-6 m = 1 m/s, cloud ice v_fall_solid of 95e-6 m =
// 0.001 m/s, cloud water v_fall_liquid of 12e-6 m; rain and graupel use the radius diag%a_rain,
// which the reference keeps at zero, so their v_fall_liquid is exactly zero
// (mo_dictionary.F90:475-527, control.F90:443).
constexpr int NCS = 5;    // condensates among them (constituents 1..5)
constexpr int GEN_K = 8;  // levels per exchange of sediment-heating terms
constexpr int BACK_K = 8;  // levels per chunk of back-substitution loads

template <class R>
__global__ void __launch_bounds__(BCOL * NTR * GEN_GROUPS) k_solve_gen(const CtxT<R> c, const StateDev old, const StateDev nw, int rk) {
  // sediment-heating terms of the condensates: csh(jl) = ((0 + a_1) - b_1) + a_2 - b_2 ...,
  // b only at the lowest layer (the surface sink term)
  __shared__ double sh_a_all[GEN_GROUPS][GEN_K][NCS][BCOL];
  __shared__ double sh_b_all[GEN_GROUPS][NCS][BCOL];
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, ncond = c.G.ncond;
  // blockDim.x is the wavefront width, so the tracer index is uniform over a wavefront; telling
  // the compiler so keeps the tracer selection in scalar registers and scalar branches
  const int lane = threadIdx.x, wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int t = wy % NTR, grp = wy / NTR;
  double(*sh_a)[NCS][BCOL] = sh_a_all[grp];
  double(*sh_b)[BCOL] = sh_b_all[grp];
  const int col_raw = (blockIdx.x * GEN_GROUPS + grp) * BCOL + lane;
  const bool active = col_raw < nx * ny;
  const int col = active ? col_raw : nx * ny - 1;  // idle lanes mirror a valid column, store nothing
  const int jk = col % nx + 1, ji = col / nx + 1;
  const double itw = 0.5, ew = 1.0 - itw;
  const double dt = c.C.dtime;
  const int jc = t < NCS ? t + 1 : 7;  // 1-based constituent index
  const bool cond = jc <= ncond;
  const bool heat = rk == 1;           // the sediment heating is diagnosed at the predictor only
  const double vf_snow = v_fall_solid(500.e-6), vf_ice = v_fall_solid(95.e-6);
  const Fld<const double> ro = old.rho[jc - 1];
  const Fld<const R> tr = c.w.tend_rho[jc - 1];
  const Fld<double> sep = c.w.col[2 * t], srp = c.w.col[2 * t + 1];
  // inputs of layer jl and of the interface below it, loaded one level ahead of their use so
  // that the loads of the next level are in flight while this level is eliminated and stored
  struct GenIn {
    double tr, wo, wn, az, Vb, ro_b, tif;
    VfIn vf;
  };
  auto load = [&](int jl) {
    GenIn g;
    g.tr = (double)tr[LG_IX(c.G, ji, jk, jl)];
    g.wo = g.wn = g.az = g.Vb = g.ro_b = g.tif = 0.0;
    g.vf = VfIn{0.0, 0.0, 0.0, 0.0, 1.0, 1.0};
    if (jl < nl) {
      const ix_t iw = LG_IX(c.G, ji, jk, jl + 1);
      g.wo = old.wind_w[iw];
      g.wn = nw.wind_w[iw];
      g.az = c.g64.area_z[iw];
      g.Vb = c.g64.volume[iw];
      g.ro_b = ro[iw];
      if (heat && cond) g.tif = (double)c.w.temperature[iw];  // temperature of layer jl+1 (both flux directions)
      if (jc == 4) g.vf = vf_load(c, old, ji, jk, jl + 1);
    }
    return g;
  };
  double ep = 0.0, rp = 0.0, cprev = 0.0, fi_up = 0.0, fr_up = 0.0, fe_up = 0.0;
  double vcw_lo = jc == 4 ? vf_liquid(vf_load(c, old, ji, jk, 1), 12.e-6, 1, nl) : 0.0;  // cloud-water fall speed
  double ro_cur = ro[LG_IX(c.G, ji, jk, 1)];
  double V = c.g64.volume[LG_IX(c.G, ji, jk, 1)];
  const double az_sfc = AT(c.g64.area_z, ji, jk, nl + 1);
  const double t_sfc = (heat && cond) ? (double)AT(c.w.temperature, ji, jk, nl) : 0.0;
  // one level of the forward sweep; cur holds the inputs of layer jl
  auto level = [&](const GenIn &cur, int jl, int j0) {
      const ix_t i = LG_IX(c.G, ji, jk, jl);
      const bool below = jl < nl;
      double fi = 0.0, fr = 0.0, fe = 0.0;
      const double vcw_layer = vcw_lo;  // v_fall of cloud water in layer jl
      if (below) {
        double vf = 0.0;
        if (jc == 1) {
          vf = 0.5 * (vf_snow + vf_snow);
        } else if (jc == 3) {
          vf = 0.5 * (vf_ice + vf_ice);
        } else if (jc == 4) {
          vcw_lo = vf_liquid(cur.vf, 12.e-6, jl + 1, nl);
          vf = 0.5 * (vcw_layer + vcw_lo);
        } else if (jc == 2 || jc == 5) {
          vf = 0.5 * (0.0 + 0.0);
        }
        fi = cur.wo - vf;
        fr = cur.wn - vf;
        fi = cur.az * fi;
        fr = cur.az * fr;
        const double dens = fr >= 0.0 ? cur.ro_b : ro_cur;
        fr = dens * fr;
        if (heat && cond) {
          const double tif = cur.tif;
          const double cp = ((jc - 1) % 2 == 0) ? c_p_ice(tif) : c_p_water(tif);
          fe = cp * tif * fr;
        }
      }
      double d, e = 0.0, cq = 0.0;
      if (jl == 1) {
        d = (fi >= 0.0) ? 1.0 : 1.0 - itw * dt / V * fi;
      } else if (jl == nl) {
        d = (fi_up >= 0.0) ? 1.0 + itw * dt / V * fi_up : 1.0;
      } else {
        d = 1.0;
        if (fi_up >= 0.0) d = d + itw * dt / V * fi_up;
        if (fi < 0.0) d = d - itw * dt / V * fi;
      }
      double vfs = 0.0;  // sink velocity in the lowest layer
      if (jl == nl && cond) {
        vfs = jc == 1 ? vf_snow : jc == 3 ? vf_ice : jc == 4 ? vcw_layer : 0.0;
        d = d + itw * vfs * dt * az_sfc / V;
      }
      if (below) {
        if (fi >= 0.0) {
          cq = 0.0;
          e = -itw * dt / V * fi;
        } else {
          cq = itw * dt / cur.Vb * fi;
          e = 0.0;
        }
      }
      double r = ro_cur + dt * cur.tr;
      double ha = 0.0, hb = 0.0;  // this tracer's sediment-heating terms of the layer
      if (jl == 1) {
        r = r + ew * dt * fr / V;
        ha = fe / V;
      } else if (jl == nl) {
        r = r - ew * dt * fr_up / V;
        ha = -(fe_up / V);
        if (cond) {
          r = r - ew * vfs * dt * ro_cur * az_sfc / V;
          if (jc > ncond / 4 && heat) hb = vfs * t_sfc * c_p_cond(jc, t_sfc) * ro_cur * az_sfc / V;
        }
      } else {
        r = r + ew * dt * (-fr_up + fr) / V;
        ha = (-fe_up + fe) / V;
      }
      if (heat && cond) {
        sh_a[jl - j0][t][lane] = ha;
        if (jl == nl) sh_b[t][lane] = hb;
      }
      double epq, rpq;
      if (jl == 1) {
        epq = e / d;
        rpq = r / d;
      } else {
        const double den = d - ep * cprev;
        epq = e / den;
        rpq = (r - rp * cprev) / den;
      }
      if (active) {
        sep[i] = epq;
        srp[i] = rpq;
      }
      ep = epq;
      rp = rpq;
      cprev = cq;
      fi_up = fi;
      fr_up = fr;
      fe_up = fe;
      ro_cur = cur.ro_b;
      V = cur.Vb;
  };
  // Two input buffers, used by odd and even levels: a buffer is refilled with the inputs two
  // levels down as soon as its level is done, so the loads of two levels are in flight.
  GenIn in_a = load(1), in_b = in_a;
  if (nl >= 2) in_b = load(2);
  for (int j0 = 1; j0 <= nl; j0 += GEN_K) {
    const int j1 = j0 + GEN_K - 1 < nl ? j0 + GEN_K - 1 : nl;
    for (int jl = j0; jl <= j1; ++jl) {
      if (jl & 1) {
        level(in_a, jl, j0);
        if (jl + 2 <= nl) in_a = load(jl + 2);
      } else {
        level(in_b, jl, j0);
        if (jl + 2 <= nl) in_b = load(jl + 2);
      }
    }
    if (heat) {  // uniform over the launch: every wave reaches both barriers
      __syncthreads();
      if (t == NTR - 1 && active) {
        for (int jl = j0; jl <= j1; ++jl) {
          double csh = 0.0;
#pragma unroll
          for (int q = 0; q < NCS; ++q) {
            csh = csh + sh_a[jl - j0][q][lane];
            if (jl == nl) csh = csh - sh_b[q][lane];
          }
          c.w.csh[LG_IX(c.G, ji, jk, jl)] = (R)csh;
        }
      }
      __syncthreads();
    }
  }
  if (!active) return;
  // back substitution and clipping of round-off negatives (mo_column_solvers.F90:584-596).
  // The scratch rows come in chunks of BACK_K levels into two register buffers: one chunk's
  // 2*BACK_K loads are in flight while the other chunk is substituted.
  double x = 0.0;
  const Fld<double> rn = nw.rho[jc - 1];
  double ar[BACK_K], ae[BACK_K], br[BACK_K], be[BACK_K];
  auto fetch = [&](int top, double *pr, double *pe) {
#pragma unroll
    for (int q = 0; q < BACK_K; ++q) {
      const int jl = top - q;
      if (jl >= 1) {
        const ix_t i = LG_IX(c.G, ji, jk, jl);
        pr[q] = srp[i];
        pe[q] = sep[i];
      }
    }
  };
  auto subst = [&](int top, const double *pr, const double *pe) {
#pragma unroll
    for (int q = 0; q < BACK_K; ++q) {
      const int jl = top - q;
      if (jl >= 1) {
        x = (jl == nl) ? pr[q] : pr[q] - pe[q] * x;
        rn[LG_IX(c.G, ji, jk, jl)] = x < 0.0 ? 0.0 : x;
      }
    }
  };
  fetch(nl, ar, ae);
  for (int top = nl; top >= 1; top -= 2 * BACK_K) {
    fetch(top - BACK_K, br, be);
    subst(top, ar, ae);
    fetch(top - 2 * BACK_K, ar, ae);
    subst(top - BACK_K, br, be);
  }
}

template __global__ void k_solve_ver<double>(const CtxT<double>, const StateDev, const StateDev, int);
template __global__ void k_solve_ver<float>(const CtxT<float>, const StateDev, const StateDev, int);
template __global__ void k_solve_gen<double>(const CtxT<double>, const StateDev, const StateDev, int);
template __global__ void k_solve_gen<float>(const CtxT<float>, const StateDev, const StateDev, int);

}  // namespace lgame
