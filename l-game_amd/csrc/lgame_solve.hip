// lgame_solve.hip -- implicit vertical column solvers of the HEVI step (FP64 in both forms, gfx950).
//
// One column per lane: consecutive lanes take consecutive x, so every per-level load and store
// of a wavefront is one coalesced 512-byte row segment.  The tridiagonal coefficients of a
// column are built while streaming down the column with a rolling window of three layers; the
// forward Thomas elimination runs in the same sweep, and only the eliminated coefficients plus
// the few per-level values the back substitution needs go to coalesced column scratch.
#include <type_traits>

#include "lgame_ops.h"
#include "lgame_step.h"
#include "lgame_batch.h"

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
  double V, rho_o, trho, rt_o, trt, ep_o, tb, eb_or_ga;  // layer Lr (exner_bg at rk 1, scratch ga at rk 2)
  double rp_u, ep_u, tin_u;                              // row Lr-1 of the system (tin from scratch at rk 2)
  double az_b, wo_b;                                     // interface Lr+1 (below the layer)
};

// row q of the interface mass-flux system (mo_column_solvers.F90:187-226): diagonal d, right-hand
// side r, and (when layer q+2 exists) the coefficients c and e that couple it to rows q+1 and q-1;
// L0, L1, L2 are layers q, q+1, q+2, F0 and F1 the interfaces q+1 and q+2
struct Row {
  double d, r, c, e;
};
__device__ __forceinline__ Row make_row(const Layer &L0, const Layer &L1, const Layer &L2, const Iface &F0, const Iface &F1,
                                        bool has_next, double dt) {
  const double K2 = IMPL_THERMO_WEIGHT * (dt * dt);
  const double K1 = IMPL_THERMO_WEIGHT * dt * C_D_P;
  Row w;
  const double dzs = L0.z - L1.z;
  const double deb = L0.eb - L1.eb;
  w.d = -F0.tin * F0.tin * (L0.ga + L1.ga) + 0.5 * deb * (L1.al - L0.al + F0.tin * (L1.be - L0.be)) -
        dzs / (K2 * C_D_P * F0.rio) * (2.0 / F0.az + dt * F0.wo * 0.5 * (-1.0 / L0.V + 1.0 / L1.V)) * F0.dp;
  w.r = -(F0.wo + dt * F0.tw) * dzs / (K2 * C_D_P) + F0.tin * (L0.epe - L1.epe) / dt + 0.5 / dt * (L0.tpe + L1.tpe) * deb -
        dzs / (K2 * C_D_P) * F0.wo * F0.rie / F0.rio;
  w.c = 0.0;
  w.e = 0.0;
  if (has_next) {
    w.c = F1.tin * L1.ga * F0.tin + 0.5 * (L1.eb - L2.eb) * (L1.al + L1.be * F0.tin) -
          (L1.z - L2.z) / K1 * 0.5 * F1.wo / (L1.V * F1.rio) * F1.dp;
    w.e = F0.tin * L1.ga * F1.tin - 0.5 * deb * (L1.al + L1.be * F1.tin) + dzs / K1 * 0.5 * F0.wo / (L1.V * F0.rio) * F0.dp;
  }
  return w;
}

// the new rho and rhotheta_v of layer Lr (mo_column_solvers.F90:307-320) from the solution at its
// upper (x_up, row Lr-1) and lower (x_low, row Lr) interface and theta_v there
__device__ __forceinline__ void new_densities(int Lr, int nl, double rhoe, double rte, double V, double dt, double x_up,
                                              double x_low, double tin_u, double tin_low, double &rho_new, double &rt_new) {
  if (Lr == nl) {
    rho_new = rhoe - dt * x_up / V;
    rt_new = rte - dt * tin_u * x_up / V;
  } else if (Lr >= 2) {
    rho_new = rhoe + dt * (-x_up + x_low) / V;
    rt_new = rte + dt * (-tin_u * x_up + tin_low * x_low) / V;
  } else {
    rho_new = rhoe + dt * x_low / V;
    rt_new = rte + dt * tin_low * x_low / V;
  }
}

// three_band_solver_ver, mo_column_solvers.F90:21-351 (soil and surface heat flux off).
//
// The back substitution needs, per layer, rho_e, rhotheta_e, gamma and, per row, rho at the
// interface and theta_v at the interface.  It forms again, with the same expressions, those it
// can from values it loads anyway (rho_old, rhotheta_old, Exner', the volume): rho_e and
// rhotheta_e from the old densities and the tendencies, the interface rho from the rolling rho of
// the layer below, and at the predictor also gamma and the interface theta_v from
// rhotheta_old/rho_old.  Only the eliminated coefficients (and at the corrector gamma and the
// interface theta_v, which need the predictor state) go through the column scratch.
template <class R, int RK>
__device__ __forceinline__ void ver_column(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int ji, int jk) {
  const int nl = c.G.nl, g = c.G.ncond;
  const StateDev &nu = RK == 1 ? old : nw;
  const double dt = c.C.dtime;
  const Fld<double> sep = c.w.col[0], srp = c.w.col[1], stin = c.w.col[4], sga = c.w.col[5];

  // forward sweep: rows q = 1..nl-1 of the interface mass-flux system
  Layer L0 = make_layer(load_raw(c, old, nu, RK, LG_IX(c.G, ji, jk, 1)), RK, dt);
  const RawLayer raw1 = load_raw(c, old, nu, RK, LG_IX(c.G, ji, jk, 2));
  RawLayer nxt = raw1;  // layer q+2 for the first row
  if (nl >= 3) nxt = load_raw(c, old, nu, RK, LG_IX(c.G, ji, jk, 3));
  Layer L1 = make_layer(raw1, RK, dt);
  Iface F0 = make_iface(L0, L1, raw1);
  Layer L2 = L1;
  Iface F1 = F0;
  double ep_prev = 0.0, rp_prev = 0.0, c_prev = 0.0;
  if (RK == 2) sga[LG_IX(c.G, ji, jk, 1)] = L0.ga;
  for (int q = 1; q <= nl - 1; ++q) {
    const bool has_next = q + 2 <= nl;
    const RawLayer cur = nxt;  // layer q+2
    if (q + 3 <= nl) nxt = load_raw(c, old, nu, RK, LG_IX(c.G, ji, jk, q + 3));
    if (RK == 2) sga[LG_IX(c.G, ji, jk, q + 1)] = L1.ga;  // layer q+1
    if (has_next) {
      L2 = make_layer(cur, RK, dt);
      F1 = make_iface(L1, L2, cur);
    }
    const Row w = make_row(L0, L1, L2, F0, F1, has_next, dt);
    const double d = w.d, r = w.r, cq = w.c, eq = w.e;
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
    if (RK == 2) stin[iq] = F0.tin;
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
    b.rho_o = old.rho[g][i];
    b.trho = (double)c.w.tend_rho[g][i];
    b.rt_o = old.rhotheta_v[i];
    b.trt = (double)c.w.tend_rhotheta_v[i];
    b.ep_o = old.exner_pert[i];
    b.tb = c.g64.theta_v_bg[i];
    b.eb_or_ga = RK == 1 ? c.g64.exner_bg[i] : sga[i];
    b.rp_u = b.ep_u = b.tin_u = 0.0;
    if (Lr >= 2) {
      const ix_t iu = LG_IX(c.G, ji, jk, Lr - 1);
      b.rp_u = srp[iu];
      b.ep_u = sep[iu];
      if (RK == 2) b.tin_u = stin[iu];
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
  bool finite = true;  // the corrector's new values, for the step's non-finite flag
  double rho_below = 0.0;    // new rho of the layer below
  double rho_o_below = 0.0;  // old rho of the layer below
  BackIn bn = load_back(nl);
  double ratio_cur = RK == 1 ? bn.rt_o / bn.rho_o : 0.0;  // rhotheta/rho of layer Lr (predictor)
  for (int Lr = nl; Lr >= 1; --Lr) {
    const BackIn b = bn;
    if (Lr >= 2) bn = load_back(Lr - 1);
    // the layer and row quantities of the forward sweep (make_layer, make_iface)
    const double rhoe = b.rho_o + dt * b.trho;
    const double rte = b.rt_o + dt * b.trt;
    const double ga = RK == 1 ? R_D / (C_D_V * b.rt_o) * (b.eb_or_ga + b.ep_o) / b.V : b.eb_or_ga;
    double tin_u = b.tin_u;  // theta_v at interface Lr
    if (RK == 1 && Lr >= 2) {
      const double ratio_up = bn.rt_o / bn.rho_o;
      tin_u = 0.5 * (ratio_up + ratio_cur);
      ratio_cur = ratio_up;
    }
    const ix_t i = LG_IX(c.G, ji, jk, Lr);
    const double V = b.V;
    double rho_new, rt_new;
    double x_up = 0.0;
    if (Lr == nl)
      x_up = b.rp_u;  // x(nl-1)
    else if (Lr >= 2)
      x_up = b.rp_u - b.ep_u * x_low;
    new_densities(Lr, nl, rhoe, rte, V, dt, x_up, x_low, tin_u, tin_low, rho_new, rt_new);
    nw.rho[g][i] = rho_new;
    nw.rhotheta_v[i] = rt_new;
    if (Lr < nl) {  // w at the interface below this layer, level Lr+1
      const double rio = 0.5 * (b.rho_o + rho_o_below);  // rho at that interface (row Lr)
      const double rin = 0.5 * (rho_new + rho_below);
      nw.wind_w[LG_IX(c.G, ji, jk, Lr + 1)] = (2.0 * x_low / b.az_b - rin * b.wo_b) / rio;
    }
    const double ep_new = b.ep_o + V * ga * (rt_new - b.rt_o);
    const double tp_new = rt_new / rho_new - b.tb;
    nw.exner_pert[i] = ep_new;
    nw.theta_v_pert[i] = tp_new;
    if (RK == 2) finite = finite && isfinite(rho_new) && isfinite(rt_new) && isfinite(ep_new) && isfinite(tp_new);
    rho_below = rho_new;
    rho_o_below = b.rho_o;
    x_low = x_up;
    tin_low = tin_u;
  }
  if (RK == 2 && !finite) atomicOr(c.C.nonfinite, 1ull);
}

template <class R>
__device__ __forceinline__ void b_solve_ver(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int rk, const Bid bid) {
  const int nx = c.G.nx, wpr = col_waves_per_row(nx);
  const int ji = bid.x / wpr + 1, jk = (bid.x % wpr) * BCOL + (int)threadIdx.x + 1;
  if (ji > c.G.ny || jk > nx) return;
  if (rk == 1)
    ver_column<R, 1>(c, old, nw, ji, jk);
  else
    ver_column<R, 2>(c, old, nw, ji, jk);
}
template <class R>
__global__ void __launch_bounds__(BCOL) k_solve_ver(const CtxT<R> c, const StateDev old, const StateDev nw, int rk) {
  b_solve_ver(c, old, nw, rk, bid_here());
}

// ---- the same solve in three passes, for domains with fewer column wavefronts than SIMDs -------
// On such a domain k_solve_ver's wavefronts leave SIMDs idle and each runs its column's chain of
// dependent FP64 operations alone.  Here the rows of the system (k_ver_rows) and the new state
// (k_ver_update) are formed by one thread per (column, level), and only the Thomas recursion
// (k_ver_thomas) runs per column.  Every value is formed by the same expressions as in ver_column
// (make_layer, make_iface, make_row, new_densities), so the results are the same bits.
// Column scratch: col[0] d, then c'; col[1] r, then r', then x; col[2] c; col[3] e (rows);
// col[4] theta_v at the row's interface (make_iface's tin); col[5] gamma of each layer.
template <class R, int RK>
__device__ __forceinline__ void ver_row(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int ji, int jk, int q) {
  const int nl = c.G.nl;
  const StateDev &nu = RK == 1 ? old : nw;
  const double dt = c.C.dtime;
  const RawLayer raw1 = load_raw(c, old, nu, RK, LG_IX(c.G, ji, jk, q + 1));
  const Layer L0 = make_layer(load_raw(c, old, nu, RK, LG_IX(c.G, ji, jk, q)), RK, dt);
  const Layer L1 = make_layer(raw1, RK, dt);
  const Iface F0 = make_iface(L0, L1, raw1);
  const bool has_next = q + 2 <= nl;
  Layer L2 = L1;
  Iface F1 = F0;
  if (has_next) {
    const RawLayer raw2 = load_raw(c, old, nu, RK, LG_IX(c.G, ji, jk, q + 2));
    L2 = make_layer(raw2, RK, dt);
    F1 = make_iface(L1, L2, raw2);
  }
  const Row w = make_row(L0, L1, L2, F0, F1, has_next, dt);
  const ix_t iq = LG_IX(c.G, ji, jk, q);
  c.w.col[0][iq] = w.d;
  c.w.col[1][iq] = w.r;
  c.w.col[2][iq] = w.c;
  c.w.col[3][iq] = w.e;
  c.w.col[4][iq] = F0.tin;
  c.w.col[5][iq] = L0.ga;
  if (q == nl - 1) c.w.col[5][LG_IX(c.G, ji, jk, nl)] = L1.ga;
}
template <class R>
__global__ void __launch_bounds__(BX * 4) k_ver_rows(const CtxT<R> c, const StateDev old, const StateDev nw, int rk) {
  const int jk = blockIdx.x * BX + threadIdx.x + 1, ji = blockIdx.y * blockDim.y + threadIdx.y + 1, q = blockIdx.z + 1;
  if (jk > c.G.nx || ji > c.G.ny || q > c.G.nl - 1) return;
  if (rk == 1)
    ver_row<R, 1>(c, old, nw, ji, jk, q);
  else
    ver_row<R, 2>(c, old, nw, ji, jk, q);
}

// thomas_algorithm (mo_column_solvers.F90:606-640) on rows 1..nl-1 of one column; the rows are
// loaded in chunks of VCH, the next chunk while the current one is eliminated
constexpr int VCH = 8;
__global__ void __launch_bounds__(BCOL) k_ver_thomas(const Geom G, Fld<double> s_d, Fld<double> s_r, Fld<const double> s_c,
                                                   Fld<const double> s_e) {
  const int wpr = col_waves_per_row(G.nx);
  const int ji = blockIdx.x / wpr + 1, jk = (blockIdx.x % wpr) * BCOL + (int)threadIdx.x + 1;
  if (ji > G.ny || jk > G.nx) return;
  const int n = G.nl - 1;
  const ix_t i0 = LG_IX(G, ji, jk, 1), lvl = (ix_t)G.lvl;
  double d[2][VCH], r[2][VCH], cc[2][VCH], e[2][VCH];
  auto load = [&](int b, int q0) {
#pragma unroll
    for (int j = 0; j < VCH; ++j)
      if (q0 + j <= n) {
        const ix_t i = i0 + (ix_t)(q0 + j - 1) * lvl;
        d[b][j] = s_d[i];
        r[b][j] = s_r[i];
        cc[b][j] = s_c[i];
        e[b][j] = s_e[i];
      }
  };
  double ep_prev = 0.0, rp_prev = 0.0, c_prev = 0.0;
  load(0, 1);
  for (int q0 = 1; q0 <= n; q0 += 2 * VCH) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int qa = q0 + h * VCH;  // the chunk in buffer h
      if (qa > n) break;
      if (qa + VCH <= n) load(1 - h, qa + VCH);
#pragma unroll
      for (int j = 0; j < VCH; ++j) {
        const int q = qa + j;
        if (q > n) break;
        double epq, rpq;
        if (q == 1) {
          epq = e[h][j] / d[h][j];
          rpq = r[h][j] / d[h][j];
        } else {
          const double den = d[h][j] - ep_prev * c_prev;
          epq = e[h][j] / den;
          rpq = (r[h][j] - rp_prev * c_prev) / den;
        }
        const ix_t i = i0 + (ix_t)(q - 1) * lvl;
        s_d[i] = epq;
        s_r[i] = rpq;
        ep_prev = epq;
        rp_prev = rpq;
        c_prev = cc[h][j];
      }
    }
  }
  // back substitution from row n up, x into s_r
  auto load_b = [&](int b, int qt) {  // rows qt, qt-1, ..., qt-VCH+1
#pragma unroll
    for (int j = 0; j < VCH; ++j)
      if (qt - j >= 1) {
        const ix_t i = i0 + (ix_t)(qt - j - 1) * lvl;
        d[b][j] = s_d[i];
        r[b][j] = s_r[i];
      }
  };
  double x_low = 0.0;
  load_b(0, n);
  for (int q0 = n; q0 >= 1; q0 -= 2 * VCH) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int qt = q0 - h * VCH;
      if (qt < 1) break;
      if (qt - VCH >= 1) load_b(1 - h, qt - VCH);
#pragma unroll
      for (int j = 0; j < VCH; ++j) {
        const int q = qt - j;
        if (q < 1) break;
        const double x = q == n ? r[h][j] : r[h][j] - d[h][j] * x_low;
        s_r[i0 + (ix_t)(q - 1) * lvl] = x;
        x_low = x;
      }
    }
  }
}

// the new state of layer Lr from the solution x (mo_column_solvers.F90:305-349)
template <class R>
__global__ void __launch_bounds__(BX * 4) k_ver_update(const CtxT<R> c, const StateDev old, const StateDev nw, int rk) {
  const int jk = blockIdx.x * BX + threadIdx.x + 1, ji = blockIdx.y * blockDim.y + threadIdx.y + 1, Lr = blockIdx.z + 1;
  const int nl = c.G.nl, g = c.G.ncond;
  if (jk > c.G.nx || ji > c.G.ny || Lr > nl) return;
  const double dt = c.C.dtime;
  const Fld<const double> sx = c.w.col[1], stin = c.w.col[4], sga = c.w.col[5];
  const ix_t i = LG_IX(c.G, ji, jk, Lr), lvl = (ix_t)c.G.lvl;
  double x_up = 0.0, x_low = 0.0, tin_u = 0.0, tin_low = 0.0;
  if (Lr >= 2) {
    x_up = sx[i - lvl];
    tin_u = stin[i - lvl];
  }
  if (Lr < nl) {
    x_low = sx[i];
    tin_low = stin[i];
  }
  const double V = c.g64.volume[i];
  const double rho_o = old.rho[g][i], rt_o = old.rhotheta_v[i];
  const double rhoe = rho_o + dt * (double)c.w.tend_rho[g][i];
  const double rte = rt_o + dt * (double)c.w.tend_rhotheta_v[i];
  double rho_new, rt_new;
  new_densities(Lr, nl, rhoe, rte, V, dt, x_up, x_low, tin_u, tin_low, rho_new, rt_new);
  nw.rho[g][i] = rho_new;
  nw.rhotheta_v[i] = rt_new;
  if (Lr < nl) {  // w at the interface below this layer, level Lr+1, with the new rho of the layer below
    const ix_t ib = i + lvl;
    const double V_b = c.g64.volume[ib];
    const double rho_o_b = old.rho[g][ib];
    const double rhoe_b = rho_o_b + dt * (double)c.w.tend_rho[g][ib];
    const double rte_b = old.rhotheta_v[ib] + dt * (double)c.w.tend_rhotheta_v[ib];
    const double x2 = Lr + 1 < nl ? sx[ib] : 0.0;
    const double tin2 = Lr + 1 < nl ? stin[ib] : 0.0;
    double rho_b, rt_b;
    new_densities(Lr + 1, nl, rhoe_b, rte_b, V_b, dt, x_low, x2, tin_low, tin2, rho_b, rt_b);
    const double rio = 0.5 * (rho_o + rho_o_b);
    const double rin = 0.5 * (rho_new + rho_b);
    nw.wind_w[ib] = (2.0 * x_low / c.g64.area_z[ib] - rin * old.wind_w[ib]) / rio;
  }
  const double ep_o = old.exner_pert[i];
  const double ep_new = ep_o + V * sga[i] * (rt_new - rt_o);
  const double tp_new = rt_new / rho_new - c.g64.theta_v_bg[i];
  nw.exner_pert[i] = ep_new;
  nw.theta_v_pert[i] = tp_new;
  if (rk == 2 && !(isfinite(rho_new) && isfinite(rt_new) && isfinite(ep_new) && isfinite(tp_new)))
    atomicOr(c.C.nonfinite, 1ull);
}

template <class R>
void launch_solve_ver(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int rk, int passes, hipStream_t st) {
  const Geom &G = c.G;
  if (passes != 3 || G.nl < 3) {
    hipLaunchKernelGGL(k_solve_ver<R>, dim3((unsigned)col_waves(G)), dim3(BCOL), 0, st, c, old, nw, rk);
    return;
  }
  const dim3 b(BX, 4), g_rows((G.nx + BX - 1) / BX, (G.ny + 3) / 4, G.nl - 1), g_up(g_rows.x, g_rows.y, G.nl);
  hipLaunchKernelGGL(k_ver_rows<R>, g_rows, b, 0, st, c, old, nw, rk);
  hipLaunchKernelGGL(k_ver_thomas, dim3((unsigned)col_waves(G)), dim3(BCOL), 0, st, G, c.w.col[0], c.w.col[1],
                     Fld<const double>(c.w.col[2]), Fld<const double>(c.w.col[3]));
  hipLaunchKernelGGL(k_ver_update<R>, g_up, b, 0, st, c, old, nw, rk);
}
template void launch_solve_ver<double>(const CtxT<double> &, const StateDev &, const StateDev &, int, int, hipStream_t);
template void launch_solve_ver<float>(const CtxT<float> &, const StateDev &, const StateDev &, int, int, hipStream_t);

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
constexpr int NCS = 5;     // condensates among them (constituents 1..5)
constexpr int GEN_K = 8;   // levels per exchange of sediment-heating terms
constexpr int GEN_RING = 2 * GEN_K;  // LDS slots of heating terms: the current and the next exchange
constexpr int BACK_K = 8;  // levels per chunk of back-substitution loads

// The five kinds of tracer, by sink velocity and heat capacity: snow and cloud ice (constant
// v_fall_solid, ice), cloud water (v_fall_liquid, water), rain (no sink, water), graupel (no sink,
// ice), vapour (no sink, not a condensate).
enum { VF_NONE = 0, VF_CONST = 1, VF_LIQUID = 2 };
enum { CP_NONE = 0, CP_ICE = 1, CP_WATER = 2 };
// position of a layer in its column
enum { POS_TOP = 0, POS_MID = 1, POS_BOTTOM = 2 };

// c_p_water with both temperature branches evaluated and the one of the reference selected, so
// that the sweep below has no divergent branch (mo_dictionary.F90:112-150)
__device__ __forceinline__ double c_p_water_sel(double temperature) {
  const double temp_c = temperature - T_0;
  const double ta = temp_c > 35.0 ? 35.0 : temp_c;
  const double d = ta - 35.0, d2 = d * d, d4 = d2 * d2;
  const double ra = 0.9979 + 3.1e-6 * d2 + 3.8e-9 * d4;
  const double tb = temp_c < -37.0 ? -37.0 : temp_c;
  const double t2 = tb * tb, t3 = t2 * tb, t4 = t2 * t2;
  const double rb = 1.000938 - 2.7052e-3 * tb - 2.3235e-5 * t2 + 4.3778e-6 * t3 + 2.7136e-7 * t4;
  return 4186.8 * (temp_c >= 0.0 ? ra : rb);
}

// inputs of layer l and of the interface below it (l+1)
struct GenIn {
  double tr, wo, wn, az, Vb, ro_b, tif, vcw;  // vcw: cloud-water sink velocity of layer l+1
};

// values carried from one layer to the next: the fluxes through the interface above, the
// density and volume of the layer and, for cloud water, its sink velocity
struct GenCarry {
  double fi_up, fr_up, fe_up, ro_cur, q, vcw_lo;  // q = itw*dt/V of the layer
};

// one row of the tridiagonal system and the layer's sediment-heating terms
struct GenRow {
  double d, e, r, cq, ha, hb;
};

// the coefficients of layer l (mo_column_solvers.F90:404-560), all selects and no branches
template <int VF, int CP, bool HEAT, int POS>
__device__ __forceinline__ GenRow gen_row(const GenIn &cur, GenCarry &s, double vf_const, double dt, double V, double az_sfc,
                                          double t_sfc, bool hb_on, int l, int nl) {
  constexpr bool cond = CP != CP_NONE;
  const double itw = 0.5, ew = 1.0 - itw;
  double fi = 0.0, fr = 0.0, fe = 0.0, q_b = 0.0;
  const double vcw_layer = s.vcw_lo;  // v_fall of cloud water in layer l
  if (POS != POS_BOTTOM) {
    double vf = 0.0;
    if (VF == VF_CONST) vf = 0.5 * (vf_const + vf_const);
    if (VF == VF_LIQUID) {
      s.vcw_lo = cur.vcw;
      vf = 0.5 * (vcw_layer + s.vcw_lo);
    }
    fi = cur.wo - vf;
    fr = cur.wn - vf;
    fi = cur.az * fi;
    fr = cur.az * fr;
    const double dens = fr >= 0.0 ? cur.ro_b : s.ro_cur;
    fr = dens * fr;
    if (HEAT && cond) {
      const double tif = cur.tif;
      const double cp = CP == CP_ICE ? c_p_ice(tif) : c_p_water_sel(tif);
      fe = cp * tif * fr;
    }
    q_b = itw * dt / cur.Vb;  // itw*dt/V of the layer below, also its q
  }
  const double q = s.q;
  GenRow w;
  if (POS == POS_TOP) {
    w.d = fi >= 0.0 ? 1.0 : 1.0 - q * fi;
  } else if (POS == POS_BOTTOM) {
    w.d = s.fi_up >= 0.0 ? 1.0 + q * s.fi_up : 1.0;
  } else {
    w.d = s.fi_up >= 0.0 ? 1.0 + q * s.fi_up : 1.0;
    w.d = fi < 0.0 ? w.d - q * fi : w.d;
  }
  double vfs = 0.0;  // sink velocity in the lowest layer
  if (POS == POS_BOTTOM && cond) {
    vfs = VF == VF_CONST ? vf_const : VF == VF_LIQUID ? vcw_layer : 0.0;
    w.d = w.d + itw * vfs * dt * az_sfc / V;
  }
  w.e = 0.0;
  w.cq = 0.0;
  if (POS != POS_BOTTOM) {
    w.e = fi >= 0.0 ? -(q * fi) : 0.0;  // -itw*dt/V*fi
    w.cq = fi >= 0.0 ? 0.0 : q_b * fi;  // itw*dt/V(below)*fi
  }
  w.r = s.ro_cur + dt * cur.tr;
  w.ha = 0.0;
  w.hb = 0.0;
  if (POS == POS_TOP) {
    w.r = w.r + ew * dt * fr / V;
    w.ha = fe / V;
  } else if (POS == POS_BOTTOM) {
    w.r = w.r - ew * dt * s.fr_up / V;
    w.ha = -(s.fe_up / V);
    if (cond) {
      w.r = w.r - ew * vfs * dt * s.ro_cur * az_sfc / V;
      if (HEAT) {
        const double cp = CP == CP_ICE ? c_p_ice(t_sfc) : c_p_water_sel(t_sfc);
        w.hb = hb_on ? vfs * t_sfc * cp * s.ro_cur * az_sfc / V : 0.0;
      }
    }
  } else {
    w.r = w.r + ew * dt * (-s.fr_up + fr) / V;
    w.ha = (-s.fe_up + fe) / V;
  }
  s.fi_up = fi;
  s.fr_up = fr;
  s.fe_up = fe;
  s.ro_cur = cur.ro_b;
  s.q = q_b;
  return w;
}

// three_band_solver_gen_densities, mo_column_solvers.F90:353-604.
//
// A block holds 64-column groups with one wavefront per tracer (snow, rain, cloud ice, cloud
// water, graupel, vapour), so every lane eliminates one tracer of one column and the per-level
// loads of the winds, metrics and temperature that all tracers of a column share are served
// from the CU's cache.  At the predictor substep the five condensates also contribute to the
// sediment heating, which the reference accumulates per level in constituent order; each
// condensate wave leaves its per-level terms in LDS, and after every GEN_K levels the vapour
// wave (which has the least work) adds them up in that order and stores the heating.
// Sink velocities: snow v_fall_solid of 500e-6 m = 1 m/s, cloud ice v_fall_solid of 95e-6 m =
// 0.001 m/s, cloud water v_fall_liquid of 12e-6 m; rain and graupel use the radius diag%a_rain,
// which the reference keeps at zero, so their v_fall_liquid is exactly zero
// (mo_dictionary.F90:475-527, control.F90:443).
//
// The sweep is software-pipelined: the row of layer l+1 is built in the same straight-line block
// as the elimination of layer l, so the long dependent chain of the elimination (two divisions
// per layer) overlaps with the independent arithmetic of the next row.  Each tracer kind runs
// its own instantiation, which has no branches on the kind.
struct GenShared {
  double a[GEN_GROUPS][GEN_RING][NCS][BCOL];
  double b[GEN_GROUPS][NCS][BCOL];
};

template <class R, int VF, int CP, bool HEAT, bool OWN_EP = true>
__device__ __forceinline__ void gen_sweep(const CtxT<R> &c, const StateDev &old, const StateDev &nw, GenShared &sh, int t, int grp,
                                          int lane, bool active, int ji, int jk, int jc, double vf_const) {
  constexpr bool cond = CP != CP_NONE;
  const int nl = c.G.nl;
  const double dt = c.C.dtime;
  const Fld<const double> ro = old.rho[jc - 1];
  const Fld<const R> tr = c.w.tend_rho[jc - 1];
  const Fld<double> sep = c.w.col[2 * t], srp = c.w.col[2 * t + 1];
  const bool is_vapour = t == NTR - 1;
  // Inputs of layer l <= nl, without branches: at l = nl the interface values come from the
  // padding level nl+1, which every field has, and are not used.
  auto load = [&](int l) {
    GenIn g;
    g.tr = (double)tr[LG_IX(c.G, ji, jk, l)];
    const ix_t iw = LG_IX(c.G, ji, jk, l + 1);
    g.wo = old.wind_w[iw];
    g.wn = nw.wind_w[iw];
    g.az = c.g64.area_z[iw];
    g.Vb = c.g64.volume[iw];
    g.ro_b = ro[iw];
    g.tif = (HEAT && cond) ? (double)c.w.temperature[iw] : 0.0;  // temperature of layer l+1
    g.vcw = VF == VF_LIQUID ? c.w.vfall_cw[iw] : 0.0;
    return g;
  };
  const double az_sfc = AT(c.g64.area_z, ji, jk, nl + 1);
  const double t_sfc = (HEAT && cond) ? (double)AT(c.w.temperature, ji, jk, nl) : 0.0;
  const bool hb_on = jc > c.G.ncond / 4;
  GenCarry s;
  s.fi_up = s.fr_up = s.fe_up = 0.0;
  s.ro_cur = ro[LG_IX(c.G, ji, jk, 1)];
  double V = c.g64.volume[LG_IX(c.G, ji, jk, 1)];
  s.q = 0.5 * dt / V;
  s.vcw_lo = VF == VF_LIQUID ? AT(c.w.vfall_cw, ji, jk, 1) : 0.0;
  // heating terms of layer l go to ring slot (l-1) % GEN_RING
  auto put_heat = [&](const GenRow &w, int l) {
    if (HEAT && cond) sh.a[grp][(l - 1) % GEN_RING][t][lane] = w.ha;
  };
  // Thomas elimination of layer l (mo_column_solvers.F90:562-582; at l = 1 ep = rp = c = 0
  // leave d and r unchanged, as the reference's first row)
  double ep = 0.0, rp = 0.0, cprev = 0.0;
  const int jk_st = active ? jk : c.G.nx + 1;  // idle lanes store into the padding column
  auto eliminate = [&](const GenRow &w, int l) {
    const double den = w.d - ep * cprev;
    const double epq = w.e / den;
    const double rpq = (w.r - rp * cprev) / den;
    const ix_t i = LG_IX(c.G, ji, jk_st, l);
    if (OWN_EP) sep[i] = epq;
    srp[i] = rpq;
    ep = epq;
    rp = rpq;
    cprev = w.cq;
  };
  // two input buffers for odd and even layers, each refilled two layers ahead once used
  GenIn in_a = load(1), in_b = load(2);
  GenRow row = gen_row<VF, CP, HEAT, POS_TOP>(in_a, s, vf_const, dt, V, az_sfc, t_sfc, hb_on, 1, nl);
  V = in_a.Vb;
  put_heat(row, 1);
  in_a = load(3);  // n_layers >= 3
  for (int j0 = 1; j0 <= nl; j0 += GEN_K) {
    const int j1 = j0 + GEN_K - 1 < nl ? j0 + GEN_K - 1 : nl;
    for (int jl = j0; jl <= j1; ++jl) {
      const int l = jl + 1;  // the layer whose row is built while layer jl is eliminated
      if (l < nl) {
        GenRow nxt;
        // (past the bottom the refill reloads layer nl, which is then never used)
        const int lf = l + 2 <= nl ? l + 2 : nl;
        if (l & 1) {
          nxt = gen_row<VF, CP, HEAT, POS_MID>(in_a, s, vf_const, dt, V, az_sfc, t_sfc, hb_on, l, nl);
          V = in_a.Vb;
          eliminate(row, jl);
          put_heat(nxt, l);
          in_a = load(lf);
        } else {
          nxt = gen_row<VF, CP, HEAT, POS_MID>(in_b, s, vf_const, dt, V, az_sfc, t_sfc, hb_on, l, nl);
          V = in_b.Vb;
          eliminate(row, jl);
          put_heat(nxt, l);
          in_b = load(lf);
        }
        row = nxt;
      } else if (l == nl) {
        const GenIn &cur = (l & 1) ? in_a : in_b;
        const GenRow nxt = gen_row<VF, CP, HEAT, POS_BOTTOM>(cur, s, vf_const, dt, V, az_sfc, t_sfc, hb_on, l, nl);
        eliminate(row, jl);
        put_heat(nxt, l);
        if (HEAT && cond) sh.b[grp][t][lane] = nxt.hb;
        row = nxt;
      } else {
        eliminate(row, jl);
      }
    }
    if (HEAT) {  // uniform over the launch: every wave reaches both barriers
      __syncthreads();
      if (is_vapour && active) {
        for (int jl = j0; jl <= j1; ++jl) {
          double csh = 0.0;
#pragma unroll
          for (int q = 0; q < NCS; ++q) {
            csh = csh + sh.a[grp][(jl - 1) % GEN_RING][q][lane];
            if (jl == nl) csh = csh - sh.b[grp][q][lane];
          }
          c.w.csh[LG_IX(c.G, ji, jk, jl)] = (R)csh;
        }
      }
      __syncthreads();
    }
  }
}

// Back substitution of one tracer.  It is the same for all tracer kinds and is called once, after
// the kind-specific sweeps, so the kernel holds one copy of its code.
__device__ __forceinline__ void gen_back(const Geom &G, Fld<const double> srp, Fld<const double> sep, Fld<double> rn, int ji, int jk,
                                         unsigned long long *nonfinite) {
  const int nl = G.nl;
  // back substitution and clipping of round-off negatives (mo_column_solvers.F90:584-596).
  // The scratch rows come in chunks of BACK_K levels into two register buffers: one chunk's
  // 2*BACK_K loads are in flight while the other chunk is substituted.  Only the chunks that
  // reach above level 1 need bounds checks.
  double x = 0.0;
  bool finite = true;
  double ar[BACK_K], ae[BACK_K], br[BACK_K], be[BACK_K];
  auto fetch = [&](auto guard, int top, double *pr, double *pe) {
#pragma unroll
    for (int q = 0; q < BACK_K; ++q) {
      const int jl = top - q;
      if (!decltype(guard)::value || jl >= 1) {
        const ix_t i = LG_IX(G, ji, jk, jl);
        pr[q] = srp[i];
        pe[q] = sep[i];
      }
    }
  };
  auto subst = [&](auto guard, int top, const double *pr, const double *pe) {
#pragma unroll
    for (int q = 0; q < BACK_K; ++q) {
      const int jl = top - q;
      if (!decltype(guard)::value || jl >= 1) {
        x = (jl == nl) ? pr[q] : pr[q] - pe[q] * x;
        rn[LG_IX(G, ji, jk, jl)] = x < 0.0 ? 0.0 : x;
        finite = finite && isfinite(x);
      }
    }
  };
  const std::true_type checked;
  const std::false_type unchecked;
  int top = nl;
  fetch(checked, top, ar, ae);
  for (; top >= 2 * BACK_K; top -= 2 * BACK_K) {
    fetch(unchecked, top - BACK_K, br, be);
    subst(unchecked, top, ar, ae);
    fetch(checked, top - 2 * BACK_K, ar, ae);
    subst(unchecked, top - BACK_K, br, be);
  }
  if (top >= 1) {
    fetch(checked, top - BACK_K, br, be);
    subst(checked, top, ar, ae);
    subst(checked, top - BACK_K, br, be);
  }
  if (nonfinite && !finite) atomicOr(nonfinite, 1ull);
}


template <class R, bool HEAT>
__device__ __forceinline__ void gen_dispatch(const CtxT<R> &c, const StateDev &old, const StateDev &nw, GenShared &sh, int t, int grp,
                                             int lane, bool active, int ji, int jk) {
  const double vf_snow = v_fall_solid(500.e-6), vf_ice = v_fall_solid(95.e-6);
  // Constituents 1..5 and vapour (7).  Without the heating, rain and graupel (no sink velocity)
  // take the vapour instantiation: their surface-sink terms are d + 0 and r - 0, which leave d
  // and r unchanged (r - (+0) = r also for r = -0; the densities are not negative).
  switch (t) {
    case 0: gen_sweep<R, VF_CONST, CP_ICE, HEAT>(c, old, nw, sh, t, grp, lane, active, ji, jk, 1, vf_snow); break;
    case 1:
      if (HEAT) gen_sweep<R, VF_NONE, CP_WATER, HEAT>(c, old, nw, sh, t, grp, lane, active, ji, jk, 2, 0.0);
      else gen_sweep<R, VF_NONE, CP_NONE, HEAT>(c, old, nw, sh, t, grp, lane, active, ji, jk, 2, 0.0);
      break;
    case 2: gen_sweep<R, VF_CONST, CP_ICE, HEAT>(c, old, nw, sh, t, grp, lane, active, ji, jk, 3, vf_ice); break;
    case 3: gen_sweep<R, VF_LIQUID, CP_WATER, HEAT>(c, old, nw, sh, t, grp, lane, active, ji, jk, 4, 0.0); break;
    case 4:
      if (HEAT) gen_sweep<R, VF_NONE, CP_ICE, HEAT, false>(c, old, nw, sh, t, grp, lane, active, ji, jk, 5, 0.0);
      else gen_sweep<R, VF_NONE, CP_NONE, HEAT, false>(c, old, nw, sh, t, grp, lane, active, ji, jk, 5, 0.0);
      break;
    default: gen_sweep<R, VF_NONE, CP_NONE, HEAT, false>(c, old, nw, sh, t, grp, lane, active, ji, jk, 7, 0.0); break;
  }
}

template <class R>
__device__ __forceinline__ void b_solve_gen(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int rk, const Bid bid) {
  // sediment-heating terms of the condensates: csh(jl) = ((0 + a_1) - b_1) + a_2 - b_2 ...,
  // b only at the lowest layer (the surface sink term)
  __shared__ GenShared sh;
  const int nx = c.G.nx, ny = c.G.ny;
  // blockDim.x is the wavefront width, so the tracer index is uniform over a wavefront; telling
  // the compiler so keeps the tracer selection in scalar registers and scalar branches
  const int lane = threadIdx.x, wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int t = wy % NTR, grp = wy / NTR;
  const int g = bid.x * GEN_GROUPS + grp, wpr = col_waves_per_row(nx);  // the group: wavefront g % wpr of row g / wpr
  const int ji_raw = g / wpr + 1, jk_raw = (g % wpr) * BCOL + lane + 1;
  const bool active = ji_raw <= ny && jk_raw <= nx;
  const int ji = active ? ji_raw : ny, jk = active ? jk_raw : nx;  // idle lanes mirror a valid column, store nothing
  if (rk == 1)  // the sediment heating is diagnosed at the predictor only
    gen_dispatch<R, true>(c, old, nw, sh, t, grp, lane, active, ji, jk);
  else
    gen_dispatch<R, false>(c, old, nw, sh, t, grp, lane, active, ji, jk);
  // Rain, graupel and vapour have no sink velocity, so their rows have the same d, e and c and
  // their eliminated coefficients e' are the same bits (the surface sink terms of rain and
  // graupel add and subtract zero); rain's sweep stores them for all three.
  __syncthreads();  // (uniform: every wavefront of the block reaches it)
  if (!active) return;
  const int jc = t < NCS ? t + 1 : 7;
  const bool shared_ep = t == 4 || t == 5;  // graupel, vapour
  gen_back(c.G, c.w.col[2 * t + 1], c.w.col[shared_ep ? 2 : 2 * t], nw.rho[jc - 1], ji, jk, rk == 2 ? c.C.nonfinite : nullptr);
}
template <class R>
__global__ void __launch_bounds__(BCOL * NTR * GEN_GROUPS) k_solve_gen(const CtxT<R> c, const StateDev old, const StateDev nw, int rk) {
  b_solve_gen(c, old, nw, rk, bid_here());
}

template __global__ void k_solve_ver<double>(const CtxT<double>, const StateDev, const StateDev, int);
template __global__ void k_solve_ver<float>(const CtxT<float>, const StateDev, const StateDev, int);
template __global__ void k_solve_gen<double>(const CtxT<double>, const StateDev, const StateDev, int);
template __global__ void k_solve_gen<float>(const CtxT<float>, const StateDev, const StateDev, int);

// ---- the same kernels over a group of domains (one launch per phase) ------------------------
LG_BATCH_KERNEL(solve_ver, BCOL, PH_SOLVE_VER, b_solve_ver(d.c, d.s[a.old_slot], d.s[a.new_slot], a.rk, bid))
LG_BATCH_KERNEL(solve_gen, BCOL * NTR * GEN_GROUPS, PH_SOLVE_GEN, b_solve_gen(d.c, d.s[a.old_slot], d.s[a.new_slot], a.rk, bid))

template <class R>
bool batch_launch_solve(int ph, dim3 g, dim3 b, hipStream_t st, const MemberDev<R> *M, const int *prefix, int nm, const BatchArgs &a) {
  if (ph == PH_SOLVE_VER) {
    hipLaunchKernelGGL(kb_solve_ver<R>, g, b, 0, st, M, prefix, nm, a);
    return true;
  }
  if (ph == PH_SOLVE_GEN) {
    hipLaunchKernelGGL(kb_solve_gen<R>, g, b, 0, st, M, prefix, nm, a);
    return true;
  }
  return false;
}
template bool batch_launch_solve<double>(int, dim3, dim3, hipStream_t, const MemberDev<double> *, const int *, int, const BatchArgs &);
template bool batch_launch_solve<float>(int, dim3, dim3, hipStream_t, const MemberDev<float> *, const int *, int, const BatchArgs &);

}  // namespace lgame
