// lgame_diff.hip -- diffusion and turbulence kernels of the HEVI step for gfx950.
//
// The first substep of vector_tend_expl (mo_vector_tend_expl.F90:101-126) and scalar_tend_expl
// (mo_scalar_tend_expl.F90:60-99) with the reference's diffusion switches on:
//   k_diff_zeta   relative vorticity zeta_z at the dual points (rel_vort, mo_vorticities.F90:19-187)
//   k_diff_coef   per cell: N^2 (update_n_squared), the TKE step (tke_update, mo_tke.F90:19-67),
//                 molecular and eddy viscosity (hor_viscosity, mo_eff_diff_coeffs.F90:24-159),
//                 the vertical coefficient molecular + tke2vert and viscosity times divergence
//   k_diff_dual   viscosity_dual times zeta_z (hor_diff_coeff_postprocessing, :161-310)
//   k_diff_mom    mom_diff_tend_x/_y (mom_diff_h, mo_momentum_diff_diss.F90:25-160)
//   k_diff_flux / k_diff_div   the diffusive fluxes and their divergence for the temperature and
//                 every constituent density (mo_scalar_tend_expl.F90:66-99)
// The vertical diffusion of w (mom_diff_v) is formed inside k_dissipation, which reads it.
// Every kernel keeps the reference's operation order (the library is built with
// -ffp-contract=off), so the FP64 form is bit-identical to the reference.
#include "lgame_ops.h"
#include "lgame_step.h"
#include "lgame_batch.h"

namespace lgame {

// square root and per-element selects of a value or a pair (each element as the one-point form)
template <class R>
__device__ __forceinline__ R vsqrt(R x) {
  return sqrt(x);
}
template <class R>
__device__ __forceinline__ P2<R> vsqrt(P2<R> x) {
  return P2<R>(sqrt(x.a), sqrt(x.b));
}
template <class F, class R>
__device__ __forceinline__ R vmap(F f, R x) {
  return f(x);
}
template <class F, class R>
__device__ __forceinline__ P2<R> vmap(F f, P2<R> x) {
  return P2<R>(f(x.a), f(x.b));
}
template <class F, class R>
__device__ __forceinline__ R vmap(F f, R x, R y) {
  return f(x, y);
}
template <class F, class R>
__device__ __forceinline__ P2<R> vmap(F f, P2<R> x, P2<R> y) {
  return P2<R>(f(x.a, y.a), f(x.b, y.b));
}

// calc_diff_coeff, mo_derived.F90:189-216: molecular kinematic viscosity of air
template <class R, class V>
__device__ __forceinline__ V calc_diff_coeff(V temperature, V density) {
  const R particle_radius = R(130e-12);
  const R particle_mass = R(M_D / N_A);
  const V thermal_velocity = vsqrt(R(8.0) * R(K_B) * temperature / (R(MPI) * particle_mass));
  const V particle_density = density / particle_mass;
  const R cross_section = R(4.0) * R(MPI) * (particle_radius * particle_radius);
  const V mean_free_path = R(1.0) / (sqrt(R(2.0)) * particle_density * cross_section);
  return R(1.0 / 3.0) * thermal_velocity * mean_free_path;
}

// tke2hor_diff_coeff, mo_eff_diff_coeffs.F90:668-683
template <class R, class V>
__device__ __forceinline__ V tke2hor(const CtxT<R> &c, V tke) {
  const V mean_velocity = vsqrt(R(2.0) * tke);
  const R mean_free_path = R(c.C.eff_hor_res) / R(6.0);
  return R(1.0 / 6.0) * mean_free_path * mean_velocity;
}

// tke2vert_diff_coeff, mo_eff_diff_coeffs.F90:685-718
template <class R, class V>
__device__ __forceinline__ V tke2vert(V tke, V n_squared, V layer_thickness) {
  const V tke_vert = R(3.0 * 1e-3) * tke;
  const V mean_velocity = vsqrt(R(2.0) * tke_vert);
  const V n_used = vsqrt(vmap([](R n) { return n > R(1e-4) ? n : R(1e-4); }, n_squared));
  V mean_free_path = vsqrt(R(2.0) * tke_vert) / n_used;
  mean_free_path = vmap([](R l, R m) { return l < m ? l : m; }, layer_thickness, mean_free_path);
  return R(1.0 / 6.0) * mean_free_path * mean_velocity;
}

// 0.5*(sum(rho(a,1:ncond+1)) + sum(rho(b,1:ncond+1))), mo_momentum_diff_diss.F90:123-159, from
// the sums k_temperature stored
template <class R>
__device__ __forceinline__ R mean_rho2(const CtxT<R> &c, ix_t a, ix_t b) {
  return R(0.5) * (c.w.rsum[a] + c.w.rsum[b]);
}

// ---------------------------------------------------------------------------------------------
// relative vorticity zeta_z = rel_vort_z_local/area_dual_z at all (ny+1) x (nx+1) dual points;
// row ny+1 and column nx+1 have their own metrics (mo_vorticities.F90:177-184).
// ---------------------------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ void b_diff_zeta(const CtxT<R> &c, const StateDev &s, const Bid bid) {
  KIDX_ZFAST(c.G.nx + 1, c.G.ny + 1, c.G.nl);
  if (jk > c.G.nx + 1 || ji > c.G.ny + 1 || jl > c.G.nl) return;
  AT(c.w.zeta, ji, jk, jl) = rel_vort_z(c, s, ji, jk, jl) / AT(c.g.area_dual_z, ji, jk, jl);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_zeta(const CtxT<R> c, const StateDev s) {
  b_diff_zeta(c, s, bid_here());
}

// ---------------------------------------------------------------------------------------------
// per-cell coefficients of the first substep
// ---------------------------------------------------------------------------------------------
// gravity_m_v * (d theta_v/dz) / theta_v at level L (2 <= L <= n_layers), the w_placeholder of
// update_n_squared (mo_eff_diff_coeffs.F90:636-648); th_a, th_b are theta_v of layers L-1, L.
template <class R, class X>
__device__ __forceinline__ VT<R, X> n2_level(const CtxT<R> &c, int ji, X jk, int L, VT<R, X> th_a, VT<R, X> th_b) {
  const VT<R, X> w = (th_a - th_b) / LDG(c.g.dz, ji, jk, L);
  const VT<R, X> inv_a = R(1.0) / th_a, inv_b = R(1.0) / th_b;
  return LDG(c.g.gravity_m_v, ji, jk, L) * (R(0.5) * (inv_a + inv_b) * w);
}

// the coefficients of one cell (jk an int) or of the columns k, k+1 (jk an XP)
template <class R, class X>
__device__ __forceinline__ void diff_coef_at(const CtxT<R> &c, const StateDev &s, int ji, X jk, int jl) {
  using V = VT<R, X>;
  const int nl = c.G.nl;
  const auto ke = east(jk);
  const V rg = LDGR(s.rho[c.G.ncond], ji, jk, jl);
  // N^2, update_n_squared (mo_eff_diff_coeffs.F90:616-666).  Its top and bottom layers take levels
  // 1 and n_levels of w_placeholder, which grad_vert does not write and which hold zero there.
  V n2 = R(0.0);
  if (jl > 1 && jl < nl) {
    const V th_m = LDG(c.g.theta_v_bg, ji, jk, jl - 1) + LDGR(s.theta_v_pert, ji, jk, jl - 1);
    const V th_c = LDG(c.g.theta_v_bg, ji, jk, jl) + LDGR(s.theta_v_pert, ji, jk, jl);
    const V th_p = LDG(c.g.theta_v_bg, ji, jk, jl + 1) + LDGR(s.theta_v_pert, ji, jk, jl + 1);
    n2 = LDG(c.g.ipw[4], ji, jk, jl) * n2_level(c, ji, jk, jl, th_m, th_c) +
         LDG(c.g.ipw[5], ji, jk, jl) * n2_level(c, ji, jk, jl + 1, th_c, th_p);
  }
  // tke_update (mo_tke.F90:19-67): advection by the inner product of the wind with grad(tke),
  // production by the dissipation of the previous step, decay with the viscosity of the previous
  // step, clipping at zero
  const Fld<const R> tk = c.w.tke;
  const V gx0 = grad_hor_x(c, tk, ji, jk, jl), gx1 = grad_hor_x(c, tk, ji, ke, jl);
  const V gy0 = grad_hor_y(c, tk, ji, jk, jl), gy1 = grad_hor_y(c, tk, ji + 1, jk, jl);
  const VGrad<R, Fld<const R>> vg{tk};
  const V gz0 = vg(c, ji, jk, jl), gz1 = vg(c, ji, jk, jl + 1);
  const V adv = LDG(c.g.ipw[0], ji, jk, jl) * LDGR(s.wind_u, ji, ke, jl) * gx1 +
                LDG(c.g.ipw[1], ji, jk, jl) * LDGR(s.wind_v, ji, jk, jl) * gy0 +
                LDG(c.g.ipw[2], ji, jk, jl) * LDGR(s.wind_u, ji, jk, jl) * gx0 +
                LDG(c.g.ipw[3], ji, jk, jl) * LDGR(s.wind_v, ji + 1, jk, jl) * gy1 +
                LDG(c.g.ipw[4], ji, jk, jl) * LDGR(s.wind_w, ji, jk, jl) * gz0 +
                LDG(c.g.ipw[5], ji, jk, jl) * LDGR(s.wind_w, ji, jk, jl + 1) * gz1;
  const V t_old = LDG(tk, ji, jk, jl);
  const V decay = R(8.0 * (MPI * MPI)) / R(c.C.mva) * LDG(c.w.visc, ji, jk, jl) / rg;
  V t = t_old + R(c.C.dtime) * (-adv + LDG(c.w.heating_diss, ji, jk, jl) / rg - decay * t_old);
  t = vmap([](R x) { return x < R(0.0) ? R(0.0) : x; }, t);
  // hor_viscosity (mo_eff_diff_coeffs.F90:24-159): molecular plus eddy viscosity, times rho
  const V mol = calc_diff_coeff<R>(LDG(c.w.temperature, ji, jk, jl), rg);
  const V div = div_h(c, Fld<const double>(s.wind_u), Fld<const double>(s.wind_v), ji, jk, jl);
  V eddy;
  if (!c.C.scheme_tke) {
    const Fld<const R> z = c.w.zeta;
    const V vc = R(0.25) * (LDG(z, ji, jk, jl) + LDG(z, ji, ke, jl) + LDG(z, ji + 1, jk, jl) + LDG(z, ji + 1, ke, jl));
    const V a = vc * vc + div * div;
    const R b = R(c.C.bg_shear) * R(c.C.bg_shear);
    eddy = R(c.C.c_s) * R(c.C.mva) * vsqrt(vmap([b](R x) { return x > b ? x : b; }, a));
  } else {
    eddy = tke2hor(c, t);
  }
  const V visc = rg * (mol + eddy);
  stv(c.w.tke_next, c.G, ji, jk, jl, t);
  stv(c.w.visc, c.G, ji, jk, jl, visc);
  stv(c.w.kv, c.G, ji, jk, jl, mol + tke2vert<R>(t, n2, LDG(c.g.layer_thickness, ji, jk, jl)));
  stv(c.w.vdiv, c.G, ji, jk, jl, visc * div);
}

template <class R>
__device__ __forceinline__ void b_diff_coef(const CtxT<R> &c, const StateDev &s, const Bid bid) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  diff_coef_at(c, s, ji, jk, jl);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_coef(const CtxT<R> c, const StateDev s) {
  b_diff_coef(c, s, bid_here());
}
// nx even: the columns k = 2*jk-1 and k+1 per thread
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_coef2(const CtxT<R> c, const StateDev s) {
  const Bid bid = bid_here();
  KIDX_ZFAST(c.G.nx / 2, c.G.ny, c.G.nl);
  const int k = 2 * jk - 1;
  if (k > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  __builtin_assume(k >= 1 && k < c.G.nx);
  diff_coef_at(c, s, ji, XP{k}, jl);
}

// ---------------------------------------------------------------------------------------------
// viscosity_dual * zeta_z at the dual points.  viscosity_dual averages the four cells around the
// dual point and is multiplied by their average gas density; each edge and corner keeps the
// reference's order of the four terms (mo_eff_diff_coeffs.F90:214-307).  viscosity_dual of row
// ny+1 and column nx+1 is that of row 1 and column 1; zeta_z there is their own.
// ---------------------------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ void b_diff_dual(const CtxT<R> &c, const StateDev &s, const Bid bid) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny;
  if (jk > nx || ji > ny || jl > c.G.nl) return;
  const int a = wrapm(ji, ny), b = ji, p = wrapm(jk, nx), q = jk;
  const Fld<const R> v = c.w.visc;
  const Fld<const double> r = s.rho[c.G.ncond];
  const R vAP = AT(v, a, p, jl), vAQ = AT(v, a, q, jl), vBP = AT(v, b, p, jl), vBQ = AT(v, b, q, jl);
  const R rAP = ATR(r, a, p, jl), rAQ = ATR(r, a, q, jl), rBP = ATR(r, b, p, jl), rBQ = ATR(r, b, q, jl);
  R vs, rs;
  if (ji == 1 && jk == 1) {
    vs = vBQ + vAQ + vBP + vAP;
    rs = rAP + rAQ + rBP + rBQ;
  } else if (ji == 1) {
    vs = vBP + vAP + vBQ + vAQ;
    rs = rBP + rAP + rBQ + rAQ;
  } else {
    vs = vAP + vAQ + vBP + vBQ;
    rs = rAP + rAQ + rBP + rBQ;
  }
  const R vd = (R(0.25) * vs) * (R(0.25) * rs);
  AT(c.w.vz, ji, jk, jl) = vd * AT(c.w.zeta, ji, jk, jl);
  if (ji == 1) AT(c.w.vz, ny + 1, jk, jl) = vd * AT(c.w.zeta, ny + 1, jk, jl);
  if (jk == 1) AT(c.w.vz, ji, nx + 1, jl) = vd * AT(c.w.zeta, ji, nx + 1, jl);
  if (ji == 1 && jk == 1) AT(c.w.vz, ny + 1, nx + 1, jl) = vd * AT(c.w.zeta, ny + 1, nx + 1, jl);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_dual(const CtxT<R> c, const StateDev s) {
  b_diff_dual(c, s, bid_here());
}

// ---------------------------------------------------------------------------------------------
// mom_diff_h (mo_momentum_diff_diss.F90:25-160): the gradient of viscosity times divergence plus
// the curl of viscosity_dual times vorticity with its terrain-following correction, divided by
// the averaged density, at the u and v points of each cell.  The terrain correction of the
// gradient reads mom_diff_tend_z as the previous step left it (grad_hor at :62 takes it as the
// vertical component).  Column nx+1 and row ny+1 are copies; the TKE of this step replaces the
// carried one.
// ---------------------------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ void b_diff_mom(const CtxT<R> &c, const StateDev &s, const Bid bid) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (jk > nx || ji > ny || jl > nl) return;
  const Geom &G = c.G;
  const ix_t i = LG_IX(G, ji, jk, jl);
  const int lu = jl > 1 ? jl - 1 : 1, ll = jl < nl ? jl + 1 : nl;
  const Fld<const R> vz = c.w.vz, zad = c.g.z_area_dual_z, vdiv = c.w.vdiv;
  auto half_dvz = [&](int a, int b) {
    return R(0.5) * (AT(vz, a, b, lu) - AT(vz, a, b, ll)) / (AT(zad, a, b, lu) - AT(zad, a, b, ll));
  };
  const VField<R, Fld<const R>> mz{c.w.mom_z};
  const R vz_c = vz[i], zad_c = zad[i], hd_c = half_dvz(ji, jk);
  {  // u point (ji, jk)
    const int km = wrapm(jk, nx);
    const R dd = c.g.dy_dual[i];
    R u = (AT(vz, ji + 1, jk, jl) - vz_c) / dd;
    const R slope = (AT(zad, ji + 1, jk, jl) - zad_c) / dd;
    u = u - slope * (half_dvz(ji + 1, jk) + hd_c);
    R gx = (vdiv[i] - AT(vdiv, ji, km, jl)) / c.g.dx[i];
    if (jl > G.nflat) gx = gx - c.g.slope_x[i] * remap_x(c, mz, ji, jk, jl);
    const R mx = (gx + u) / mean_rho2(c, LG_IX(G, ji, km, jl), i);
    c.w.mom_x[i] = mx;
    if (jk == 1) AT(c.w.mom_x, ji, nx + 1, jl) = mx;
  }
  {  // v point (ji, jk)
    const int jm = wrapm(ji, ny);
    const R dd = c.g.dx_dual[i];
    R v = (AT(vz, ji, jk + 1, jl) - vz_c) / dd;
    const R slope = (AT(zad, ji, jk + 1, jl) - zad_c) / dd;
    v = v - slope * (half_dvz(ji, jk + 1) + hd_c);
    R gy = (AT(vdiv, jm, jk, jl) - vdiv[i]) / c.g.dy[i];
    if (jl > G.nflat) gy = gy - c.g.slope_y[i] * remap_y(c, mz, ji, jk, jl);
    const R my = (gy + v) / mean_rho2(c, LG_IX(G, jm, jk, jl), i);
    c.w.mom_y[i] = my;
    if (ji == 1) AT(c.w.mom_y, ny + 1, jk, jl) = my;
  }
  c.w.tke[i] = c.w.tke_next[i];
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_mom(const CtxT<R> c, const StateDev s) {
  b_diff_mom(c, s, bid_here());
}

// ---------------------------------------------------------------------------------------------
// scalar diffusion (mo_scalar_tend_expl.F90:66-99 with scalar_diff_coeffs,
// mo_eff_diff_coeffs.F90:526-614) of the temperature (field 0) and of every constituent density
// (fields 1..nc), in two passes over the cells that each treat all fields: the grid metrics and
// the coefficients are loaded once per cell.
//   k_diff_flux: flux coefficient times grad_hor (with its terrain correction) at the u and v
//                points of the cell and times grad_vert at its upper w level
//                (scalar_times_vector_h/_v, mo_multiplications.F90:18-66,289-312);
//   k_diff_div:  div_h of the horizontal fluxes plus add_vertical_div of the vertical one
//                (mo_divergence_operators.F90:16-68,149-186).
// Coefficients: temperature c_g_v*viscosity and rho_gas*c_g_v*kv, densities viscosity/rho_gas
// and kv; c_g_v is spec_heat_cap_diagnostics_v (mo_derived.F90:40-68) with its one relevant
// constituent.  Column nx+1 and row ny+1 of the horizontal fluxes are copies of column 1 and
// row 1 (scalar_times_vector_h stores them so).
// ---------------------------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ R c_g_v(R rg) {
  R r = R(0.0);
  r = r + rg / rg * R(C_D_V);
  return r;
}

template <class R>
__device__ __forceinline__ P2<R> c_g_v(P2<R> rg) {
  return P2<R>(c_g_v(rg.a), c_g_v(rg.b));
}

// the fluxes of one cell (jk an int) or of the columns k, k+1 (jk an XP)
template <class R, class X>
__device__ __forceinline__ void diff_flux_at(const CtxT<R> &c, const StateDev &s, int ji, X jk, int jl) {
  using V = VT<R, X>;
  const Geom &G = c.G;
  const int nx = G.nx, ny = G.ny, nl = G.nl;
  const auto km = wrapm(jk, nx);
  const int jm = wrapm(ji, ny);
  const bool terr = jl > G.nflat;                 // grad_hor's terrain correction
  const bool below = jl < nl || c.C.lfreeslip;    // remap_ver2hor reads level jl+1
  const bool up = jl >= 2, dn = jl < nl;          // grad_vert is nonzero at level jl / jl+1
  // metrics
  const V dx = LDG(c.g.dx, ji, jk, jl), dy = LDG(c.g.dy, ji, jk, jl);
  V sx = R(0.0), sy = R(0.0), p4i = R(0.0), p4w = R(0.0), p4s = R(0.0), p5i = R(0.0), p5w = R(0.0), p5s = R(0.0);
  V dzi = R(1.0), dzw = R(1.0), dzs = R(1.0), dzi1 = R(1.0), dzw1 = R(1.0), dzs1 = R(1.0);
  if (up) dzi = LDG(c.g.dz, ji, jk, jl);
  if (terr) {
    sx = LDG(c.g.slope_x, ji, jk, jl);
    sy = LDG(c.g.slope_y, ji, jk, jl);
    p4i = LDG(c.g.ipw[4], ji, jk, jl);
    p4w = LDG(c.g.ipw[4], ji, km, jl);
    p4s = LDG(c.g.ipw[4], jm, jk, jl);
    if (up) {
      dzw = LDG(c.g.dz, ji, km, jl);
      dzs = LDG(c.g.dz, jm, jk, jl);
    }
    if (below) {
      p5i = LDG(c.g.ipw[5], ji, jk, jl);
      p5w = LDG(c.g.ipw[5], ji, km, jl);
      p5s = LDG(c.g.ipw[5], jm, jk, jl);
    }
    if (dn) {
      dzi1 = LDG(c.g.dz, ji, jk, jl + 1);
      dzw1 = LDG(c.g.dz, ji, km, jl + 1);
      dzs1 = LDG(c.g.dz, jm, jk, jl + 1);
    }
  }
  // coefficients at the cell, its west and south neighbours and the cell above
  const Fld<const double> rgf = s.rho[G.ncond];
  const V rgi = LDGR(rgf, ji, jk, jl), rgw = LDGR(rgf, ji, km, jl), rgs = LDGR(rgf, jm, jk, jl);
  const V vi = LDG(c.w.visc, ji, jk, jl), vw = LDG(c.w.visc, ji, km, jl), vs = LDG(c.w.visc, jm, jk, jl);
  const V kvi = LDG(c.w.kv, ji, jk, jl);
  V kva = R(0.0), rga = R(1.0);
  if (up) {
    kva = LDG(c.w.kv, ji, jk, jl - 1);
    rga = LDGR(rgf, ji, jk, jl - 1);
  }
  // one field: gradients at the u, v and w points, times the averaged coefficients
  auto one = [&](auto f, Fld<R> du, Fld<R> dv, Fld<R> dw, V chu, V chv, V cvw, bool lvert) {
    const V fi = LDGR(f, ji, jk, jl), fw = LDGR(f, ji, km, jl), fs = LDGR(f, jm, jk, jl);
    V vgi0 = R(0.0), vgw0 = R(0.0), vgs0 = R(0.0), vgi1 = R(0.0), vgw1 = R(0.0), vgs1 = R(0.0);
    if (up) vgi0 = (LDGR(f, ji, jk, jl - 1) - fi) / dzi;
    V gx = (fi - fw) / dx, gy = (fs - fi) / dy;
    if (terr) {
      if (up) {
        vgw0 = (LDGR(f, ji, km, jl - 1) - fw) / dzw;
        vgs0 = (LDGR(f, jm, jk, jl - 1) - fs) / dzs;
      }
      if (dn) {
        vgi1 = (fi - LDGR(f, ji, jk, jl + 1)) / dzi1;
        vgw1 = (fw - LDGR(f, ji, km, jl + 1)) / dzw1;
        vgs1 = (fs - LDGR(f, jm, jk, jl + 1)) / dzs1;
      }
      V r = R(0.0);  // remap_ver2hor_x, mo_averaging.F90:56-95
      r = r + p4w * vgw0;
      r = r + p4i * vgi0;
      if (below) {
        r = r + p5w * vgw1;
        r = r + p5i * vgi1;
      }
      gx = gx - sx * (R(0.5) * r);
      r = R(0.0);  // remap_ver2hor_y, :97-136
      r = r + p4s * vgs0;
      r = r + p4i * vgi0;
      if (below) {
        r = r + p5s * vgs1;
        r = r + p5i * vgi1;
      }
      gy = gy - sy * (R(0.5) * r);
    }
    const V fu = chu * gx, fv = chv * gy;
    stv(du, G, ji, jk, jl, fu);
    if (col1(jk)) AT(du, ji, nx + 1, jl) = first(fu);
    stv(dv, G, ji, jk, jl, fv);
    if (ji == 1) stv(dv, G, ny + 1, jk, jl, fv);
    if (lvert && up) stv(dw, G, ji, jk, jl, cvw * vgi0);
  };
  if (c.C.ltemp_h) {
    const V chu = R(0.5) * (c_g_v(rgw) * vw + c_g_v(rgi) * vi);
    const V chv = R(0.5) * (c_g_v(rgs) * vs + c_g_v(rgi) * vi);
    const V cvw = R(0.5) * (rga * c_g_v(rga) * kva + rgi * c_g_v(rgi) * kvi);
    one(Fld<const R>(c.w.temperature), c.w.dfu[0], c.w.dfv[0], c.w.dfw[0], chu, chv, cvw, c.C.ltemp_v != 0);
  }
  if (c.C.lmass_h) {
    const V chu = R(0.5) * (vw / rgw + vi / rgi);
    const V chv = R(0.5) * (vs / rgs + vi / rgi);
    const V cvw = R(0.5) * (kva + kvi);
    for (int jc = 0; jc < G.nc; ++jc)
      one(Fld<const double>(s.rho[jc]), c.w.dfu[jc + 1], c.w.dfv[jc + 1], c.w.dfw[jc + 1], chu, chv, cvw, c.C.lmass_v != 0);
  }
}

template <class R>
__device__ __forceinline__ void b_diff_flux(const CtxT<R> &c, const StateDev &s, const Bid bid) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  diff_flux_at(c, s, ji, jk, jl);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_flux(const CtxT<R> c, const StateDev s) {
  b_diff_flux(c, s, bid_here());
}
// nx even: the columns k = 2*jk-1 and k+1 per thread
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_flux2(const CtxT<R> c, const StateDev s) {
  const Bid bid = bid_here();
  KIDX_ZFAST(c.G.nx / 2, c.G.ny, c.G.nl);
  const int k = 2 * jk - 1;
  if (k > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  __builtin_assume(k >= 1 && k < c.G.nx);
  diff_flux_at(c, s, ji, XP{k}, jl);
}

// the divergences of one cell (jk an int) or of the columns k, k+1 (jk an XP)
template <class R, class X>
__device__ __forceinline__ void diff_div_at(const CtxT<R> &c, int ji, X jk, int jl) {
  using V = VT<R, X>;
  const Geom &G = c.G;
  const int nl = G.nl, nflat = G.nflat;
  const auto ke = east(jk);
  // which interfaces carry a contravariant correction, and which layers those need (as in
  // the scalar divergence): vcc at interface L >= nflat+1 uses layer L, and layer L-1 unless L = nflat+1
  const bool up = jl > nflat;
  const bool lo = jl >= nflat && jl < nl;
  const bool lay_m = up && jl != nflat + 1;
  const bool lay_c = up || (lo && jl != nflat);
  VccGeo<V> gm{}, gc{}, gp{};
  if (lay_m) gm = vcc_geo(c, ji, jk, jl - 1);
  if (lay_c) gc = vcc_geo(c, ji, jk, jl);
  if (lo) gp = vcc_geo(c, ji, jk, jl + 1);
  const V ax1 = LDG(c.g.area_x, ji, ke, jl), ay0 = LDG(c.g.area_y, ji, jk, jl), ax0 = LDG(c.g.area_x, ji, jk, jl),
          ay1 = LDG(c.g.area_y, ji + 1, jk, jl);
  const V az0 = LDG(c.g.area_z, ji, jk, jl), az1 = LDG(c.g.area_z, ji, jk, jl + 1), Vol = LDG(c.g.volume, ji, jk, jl);
  auto one = [&](int q, Fld<R> out, bool lvert) {
    const Fld<const R> du = c.w.dfu[q], dv = c.w.dfv[q];
    const Faces<V> fc = faces<R>(du, dv, G, ji, jk, jl);
    V cu = R(0.0), cl = R(0.0);
    if (up) {
      if (lay_m) cu = vcc_sub(cu, gm, faces<R>(du, dv, G, ji, jk, jl - 1));
      cu = vcc_sub(cu, gc, fc);
    }
    if (lo) {
      if (jl != nflat) cl = vcc_sub(cl, gc, fc);
      cl = vcc_sub(cl, gp, faces<R>(du, dv, G, ji, jk, jl + 1));
    }
    const V comp_h = fc.e * ax1 + fc.n * ay0 - fc.w * ax0 - fc.s * ay1;
    V comp_v = R(0.0);
    if (jl == nflat) {
      comp_v = -cl * az1;
    } else if (jl == nl) {
      comp_v = cu * az0;
    } else if (jl > nflat) {
      comp_v = cu * az0 - cl * az1;
    }
    V d = (comp_h + comp_v) / Vol;
    if (lvert) {
      V wu = R(0.0), wl = R(0.0);
      if (jl != 1) wu = LDG(c.w.dfw[q], ji, jk, jl);
      if (jl != nl) wl = LDG(c.w.dfw[q], ji, jk, jl + 1);
      d = d + (wu * az0 - wl * az1) / Vol;
    }
    stv(out, G, ji, jk, jl, d);
  };
  if (c.C.ltemp_h) one(0, c.w.tdh, c.C.ltemp_v != 0);
  if (c.C.lmass_h)
    for (int jc = 0; jc < G.nc; ++jc) one(jc + 1, c.w.mdt[jc], c.C.lmass_v != 0);
}

template <class R>
__device__ __forceinline__ void b_diff_div(const CtxT<R> &c, const Bid bid) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  diff_div_at(c, ji, jk, jl);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_div(const CtxT<R> c) {
  b_diff_div(c, bid_here());
}
// nx even: the columns k = 2*jk-1 and k+1 per thread
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_div2(const CtxT<R> c) {
  const Bid bid = bid_here();
  KIDX_ZFAST(c.G.nx / 2, c.G.ny, c.G.nl);
  const int k = 2 * jk - 1;
  if (k > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  __builtin_assume(k >= 1 && k < c.G.nx);
  diff_div_at(c, ji, XP{k}, jl);
}

// ---------------------------------------------------------------------------------------------
// planetary boundary layer (mo_pbl.F90): surface turbulence quantities of each column at the
// start of the step, and the friction of the lowest-layer wind in the first substep.  sfc holds
// the carried diag%roughness_length (level 1), roughness_velocity (2) and monin_obukhov_length (3)
// as 2-D fields.  The transcendental functions are the device's, so these results agree with the
// reference to rounding, not bit for bit.
// ---------------------------------------------------------------------------------------------
constexpr double EPSILON_SECURITY = 1e-10;  // mo_constants.F90:34
constexpr double GRAVITY = 9.80616;         // mo_constants.F90:21

template <class R>
__device__ __forceinline__ R fmax2(R a, R b) { return a > b ? a : b; }
template <class R>
__device__ __forceinline__ R fmin2(R a, R b) { return a < b ? a : b; }

// roughness_length_from_u10_sea, mo_pbl.F90:252-272
template <class R>
__device__ R roughness_length_from_u10_sea(R u10) {
  const R swh = R(0.0248) * (u10 * u10);
  const R period = R(0.729) * u10;
  const R wavelength = R(GRAVITY) * (period * period) / R(2.0 * MPI);
  const R r = R(1200.0) * swh * pow(swh / fmax2(wavelength, R(EPSILON_SECURITY)), R(4.5));
  return fmax2(R(0.0001), r);
}

// roughness_velocity, mo_pbl.F90:344-366
template <class R>
__device__ R roughness_velocity(const CtxT<R> &c, R wind_speed, R z_agl, R z0) {
  R den = log(z_agl / z0);
  if (fabs(den) < R(EPSILON_SECURITY)) den = R(EPSILON_SECURITY);
  return fmax2(R(EPSILON_SECURITY), wind_speed * R(c.C.karman) / den);
}

// psi_m, mo_pbl.F90:392-420
template <class R>
__device__ R psi_m(R eff, R l) {
  R ll = l;
  if (fabs(ll) < R(EPSILON_SECURITY)) ll = R(EPSILON_SECURITY);
  if (ll < R(0.0)) {
    const R x = pow(R(1.0) - R(15.0) * eff / ll, R(0.25));
    return R(2.0) * log((R(1.0) + x) / R(2.0)) + log((R(1.0) + x * x) / R(2.0)) - R(2.0) * atan(x) + R(MPI / 2.0);
  }
  return -(R(4.7) * eff / ll);
}

// the PBL friction at a lowest-layer u or v point between cells a and b (mo_pbl.F90:125-160,
// with momentum_flux_resistance, :318-342)
template <class R>
__device__ R pbl_friction(const CtxT<R> &c, R wind, R z_pt, ix_t a, ix_t b, ix_t as, ix_t bs, ix_t dl) {
  const Fld<const R> vsq = c.w.v_squared, zw = c.g.z_w, sfc = c.w.sfc;
  const R ws = R(0.5) * (sqrt(vsq[a]) + sqrt(vsq[b]));
  const R z_agl = z_pt - R(0.5) * (zw[a + dl] + zw[b + dl]);
  const R lt = R(0.5) * (zw[a] + zw[b]) - R(0.5) * (zw[a + dl] + zw[b + dl]);
  const R z0 = R(0.5) * (sfc[as] + sfc[bs]);
  const R l = R(0.5) * (sfc[as + 2 * dl] + sfc[bs + 2 * dl]);
  const R hp = R(c.C.h_prandtl), dt = R(c.C.dtime);
  const R used = fmin2(z_agl, hp);
  R fr = R(1.0) / (R(c.C.karman) * roughness_velocity(c, ws, z_agl, z0)) * (log(used / z0) - psi_m(used, l));
  if (fr < dt / z_agl) fr = dt / z_agl;
  R f = R(1.0);
  if (z_agl > hp) f = log(hp / z0) / log(z_agl / z0);
  return f * wind / fr / lt;
}

// psi_h, mo_pbl.F90:368-390
template <class R>
__device__ R psi_h(R eff, R l) {
  R ll = l;
  if (fabs(ll) < R(EPSILON_SECURITY)) ll = R(EPSILON_SECURITY);
  if (ll < R(0.0)) {
    const R x = pow(R(1.0) - R(15.0) * eff / ll, R(0.25));
    return R(2.0) * log((R(1.0) + x * x) / R(2.0));
  }
  return -(R(4.0) * eff / ll);
}

// scalar_flux_resistance, mo_pbl.F90:294-316
template <class R>
__device__ R scalar_flux_resistance(const CtxT<R> &c, R us, R z_agl, R z0, R l) {
  const R used = fmin2(z_agl, R(c.C.h_prandtl));
  R r = R(1.0) / (R(c.C.karman) * us) * (log(used / z0) - psi_h(used, l) + log(R(7.0)));
  if (r < R(c.C.dtime) / z_agl) r = R(c.C.dtime) / z_agl;
  return r;
}

// update_sfc_turb_quantities, mo_pbl.F90:19-109, one column per thread, with the scalar flux
// resistance of the prognostic soil; v_squared is the kinetic energy the last momentum advection
// block left
template <class R>
__device__ __forceinline__ void b_sfc_turb(const CtxT<R> &c, const StateDev &s, const Bid bid) {
  KIDX;
  const Geom &G = c.G;
  const int nl = G.nl;
  if (jk > G.nx || ji > G.ny || jl > 1) return;
  const ix_t dl = (ix_t)G.lvl;
  const ix_t i1 = LG_IX(G, ji, jk, 1), il = LG_IX(G, ji, jk, nl);
  const Fld<R> sfc = c.w.sfc;
  const R agl = c.g.z_scalar[il] - c.g.z_w[il + dl];
  const R u_low = sqrt(c.w.v_squared[il]);
  R z0 = sfc[i1];
  const R u10 = u_low * log(R(10.0) / z0) / log(agl / z0);
  const R lf = c.g.land_fraction[i1];
  if (lf < R(1.0)) {
    z0 = lf * c.g.roughness_length[i1] + (R(1.0) - lf) * roughness_length_from_u10_sea(u10);
    sfc[i1] = z0;
  }
  const R us = roughness_velocity(c, u_low, agl, z0);
  sfc[i1 + dl] = us;
  const R th_l = c.g.theta_v_bg[il] + R(s.theta_v_pert[il]);
  const R th_2 = c.g.theta_v_bg[il - dl] + R(s.theta_v_pert[il - dl]);
  const R dz = c.g.z_scalar[il - dl] - c.g.z_scalar[il];
  const R dth = (th_2 - th_l) / dz;
  const R w_pert = u10 * fmax2(R(0.001), R(0.02) * (R(1.0) - dth / R(0.01)));
  const R t_pert = -(R(0.2) * R(c.C.dtime) * w_pert * dth);
  R wt = R(0.2) * w_pert * t_pert;
  if (fabs(wt) < R(EPSILON_SECURITY)) wt = R(EPSILON_SECURITY);
  const R mo_len = -(th_l * (us * (us * us)) / (R(c.C.karman) * R(GRAVITY) * wt));
  sfc[i1 + 2 * dl] = mo_len;
  if (c.C.lsoil) c.w.sfc2[i1] = scalar_flux_resistance(c, us, agl, z0, mo_len);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_sfc_turb(const CtxT<R> c, const StateDev s) {
  b_sfc_turb(c, s, bid_here());
}

// pbl_wind_tendency, mo_pbl.F90:111-250: mom_diff_tend_x/_y of the lowest layer minus the
// friction, at the u and v points of each bottom cell; column nx+1 and row ny+1 are copies
template <class R>
__device__ __forceinline__ void b_pbl_wind(const CtxT<R> &c, const StateDev &s, const Bid bid) {
  KIDX;
  const Geom &G = c.G;
  const int nx = G.nx, ny = G.ny, nl = G.nl;
  if (jk > nx || ji > ny || jl > 1) return;
  const ix_t dl = (ix_t)G.lvl;
  const int km = wrapm(jk, nx), jm = wrapm(ji, ny);
  const ix_t i = LG_IX(G, ji, jk, nl), iw = LG_IX(G, ji, km, nl), is = LG_IX(G, jm, jk, nl);
  const ix_t i1 = LG_IX(G, ji, jk, 1), iw1 = LG_IX(G, ji, km, 1), is1 = LG_IX(G, jm, jk, 1);
  const R mx = c.w.mom_x[i] - pbl_friction(c, R(s.wind_u[i]), c.g.z_u[i], iw, i, iw1, i1, dl);
  const R my = c.w.mom_y[i] - pbl_friction(c, R(s.wind_v[i]), c.g.z_v[i], is, i, is1, i1, dl);
  c.w.mom_x[i] = mx;
  if (jk == 1) AT(c.w.mom_x, ji, nx + 1, nl) = mx;
  c.w.mom_y[i] = my;
  if (ji == 1) AT(c.w.mom_y, ny + 1, jk, nl) = my;
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_pbl_wind(const CtxT<R> c, const StateDev s) {
  b_pbl_wind(c, s, bid_here());
}

// ---------------------------------------------------------------------------------------------
// the surface of three_band_solver_ver (mo_column_solvers.F90:21-351), one column per thread, in
// FP64 in both forms: the sensible power flux densities over soil and sea and their heating of
// the lowest layer (:81-120), and the soil temperature rows of the column's linear system
// (:202-204, :229-300, :331-336).  In the reference's combined system the lower and upper
// diagonals vanish at row n_layers-1 (c_vector(n_layers-1) = e_vector(n_layers-1) = 0), so the
// elimination of the soil rows neither reads nor changes anything of the atmospheric rows, and
// eliminating the soil rows on their own, as here, gives the same values.  The short- and
// longwave surface terms are those of the last radiation update, zero without radiation.
// ---------------------------------------------------------------------------------------------

// the fields of a state that the sensible heat flux reads
struct SfcState {
  Fld<const double> ep, tvp, rho_g, rho_v, ts;
};

// the gas temperature of the lowest layer (mo_column_solvers.F90:84-93)
__device__ __forceinline__ double t_gas_lowest(const GridT<double> &g, const SfcState &s, bool moist, ix_t il) {
  double t = (g.exner_bg[il] + s.ep[il]) * (g.theta_v_bg[il] + s.tvp[il]);
  if (moist) t = t / (1.0 + s.rho_v[il] / s.rho_g[il] * (M_D / M_V - 1.0));
  return t;
}

template <class R>
__device__ __forceinline__ void b_surface(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int rk, const Bid bid) {
  KIDX;
  const Geom &G = c.G;
  const int nl = G.nl;
  if (jk > G.nx || ji > G.ny || jl > 1) return;
  // the state fields read, picked one by one (a reference to either state argument would place
  // both in scratch memory)
  const bool moist = c.C.lmoist != 0;
  const int gi = G.ncond, vi = moist ? G.ncond + 1 : G.ncond;
  const SfcState so{old.exner_pert, old.theta_v_pert, old.rho[gi], old.rho[vi], old.temperature_soil};
  const SfcState sn = rk == 1 ? so : SfcState{nw.exner_pert, nw.theta_v_pert, nw.rho[gi], nw.rho[vi], nw.temperature_soil};
  const GridT<double> &g = c.g64;
  const ix_t dl = (ix_t)G.lvl;
  const ix_t i1 = LG_IX(G, ji, jk, 1), il = LG_IX(G, ji, jk, nl);
  const double dt = c.C.dtime;
  double ps = 0.0;  // diag%power_flux_density_sens_soil, zero without the sensible heat flux
  if (c.C.lsens) {
    const double t_old = t_gas_lowest(g, so, moist, il);
    const double t_new = t_gas_lowest(g, sn, moist, il);
    const double lf = g.land_fraction[i1], kf = g.soil[i1], sst = g.soil[i1 + 4 * dl];
    const double rn = sn.rho_g[il], ro = so.rho_g[il];
    const double sfr = (double)c.w.sfc2[i1];
    ps = (lf + kf) * 0.5 * C_D_V * (rn * (t_old - so.ts[i1]) + ro * (t_new - sn.ts[i1])) / sfr;
    const double pe = (1.0 - lf - kf) * 0.5 * C_D_V * (rn * (t_old - sst) + ro * (t_new - sst)) / sfr;
    c.w.sfc2[i1 + dl] = (R)ps;
    c.w.sfc2[i1 + 2 * dl] = (R)pe;
    const double trt = (double)c.w.tend_rhotheta_v[il];
    c.w.tend_rhotheta_v[il] = (R)(trt - g.area_z[il + dl] * (ps + pe) / ((g.exner_bg[il] + sn.ep[il]) * C_D_P) / g.volume[il]);
  }
  if (!c.C.lsoil) return;
  // the soil rows: heat conduction between the layers and to the constant temperature at
  // z_t_const below, the surface fluxes into the top layer, Crank-Nicolson in time
  const int ns = c.C.nsoil;
  const double *zi = g.soil_z, *zc = zi + ns + 1;
  const double z_t_const = zc[ns];
  const double rc = g.soil[i1 + dl], tc = g.soil[i1 + 2 * dl], tcs = g.soil[i1 + 3 * dl];
  // diag%sfc_sw_in and sfc_lw_out of the last radiation update (zero without radiation); the
  // latent heat flux of lakes is zero (nothing in the reference calls calc_h2otracers_source_rates,
  // mo_phase_trans.F90:22-330, where it is set)
  const int col = (ji - 1) * G.nx + (jk - 1);
  const double sw_in = c.C.rad_sw_in ? c.C.rad_sw_in[col] : 0.0;
  const double lw_out = c.C.rad_lw_out ? c.C.rad_lw_out[col] : 0.0;
  const double lat_lake = 0.0;
  // rows (:229-288) and the forward sweep of thomas_algorithm (:625-634), one row at a time: the
  // explicit heat flux densities hfd(j) of the layer above and of this layer, and the soil
  // temperatures of this layer and the one below, roll along
  double ep[MAX_SOIL], rp[MAX_SOIL];
  double t_cur = so.ts[i1];
  double h_up = 0.0;    // hfd of the layer above
  double c_prev = 0.0;  // c_vector of the row above (the lower diagonal entry of this row)
#pragma unroll
  for (int j = 0; j < MAX_SOIL; ++j) {
    if (j < ns) {
      double t_next = 0.0, h;
      if (j < ns - 1) {
        t_next = so.ts[i1 + (ix_t)(j + 1) * dl];
        h = -rc * tc * (t_cur - t_next) / (zc[j] - zc[j + 1]);
      } else {
        h = -rc * tc * (t_cur - tcs) / (2.0 * (zc[j] - z_t_const));
      }
      double r;
      if (j == 0)
        r = t_cur + (ps + lat_lake + sw_in - lw_out + 0.5 * h) / ((zi[0] - zi[1]) * rc) * dt;
      else
        r = t_cur + 0.5 * (-h_up + h) / ((zi[j] - zi[j + 1]) * rc) * dt;
      const double f = 0.5 * dt * rc * tc / ((zi[j] - zi[j + 1]) * rc);
      double d;
      if (j == 0)
        d = 1.0 + f * 1.0 / (zc[j] - zc[j + 1]);
      else if (j == ns - 1)
        d = 1.0 + f * 1.0 / (zc[j - 1] - zc[j]);
      else
        d = 1.0 + f * (1.0 / (zc[j - 1] - zc[j]) + 1.0 / (zc[j] - zc[j + 1]));
      double e = 0.0, cj = 0.0;
      if (j < ns - 1) {
        cj = -0.5 * dt * rc * tc / ((zi[j + 1] - zi[j + 2]) * rc) / (zc[j] - zc[j + 1]);
        e = -0.5 * dt * rc * tc / ((zi[j] - zi[j + 1]) * rc) / (zc[j] - zc[j + 1]);
      }
      if (j == 0) {
        ep[0] = e / d;
        rp[0] = r / d;
      } else {
        if (j < ns - 1) ep[j] = e / (d - ep[j - 1] * c_prev);
        rp[j] = (r - rp[j - 1] * c_prev) / (d - ep[j - 1] * c_prev);
      }
      c_prev = cj;
      h_up = h;
      t_cur = t_next;
    }
  }
  // upward sweep (:636-639) and the new soil temperature (:331-336)
  double x = 0.0;
#pragma unroll
  for (int j = MAX_SOIL - 1; j >= 0; --j) {
    if (j < ns) {
      x = j == ns - 1 ? rp[j] : rp[j] - ep[j] * x;
      nw.temperature_soil[i1 + (ix_t)j * dl] = x;
    }
  }
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_surface(const CtxT<R> c, const StateDev old, const StateDev nw, int rk) {
  b_surface(c, old, nw, rk, bid_here());
}

#define LG_INSTANTIATE(R)                                                                  \
  template __global__ void k_diff_zeta<R>(const CtxT<R>, const StateDev);                  \
  template __global__ void k_diff_coef<R>(const CtxT<R>, const StateDev);                  \
  template __global__ void k_diff_dual<R>(const CtxT<R>, const StateDev);                  \
  template __global__ void k_diff_mom<R>(const CtxT<R>, const StateDev);                   \
  template __global__ void k_diff_flux<R>(const CtxT<R>, const StateDev);                  \
  template __global__ void k_diff_flux2<R>(const CtxT<R>, const StateDev);                 \
  template __global__ void k_diff_div<R>(const CtxT<R>);                                   \
  template __global__ void k_sfc_turb<R>(const CtxT<R>, const StateDev);                   \
  template __global__ void k_pbl_wind<R>(const CtxT<R>, const StateDev);                   \
  template __global__ void k_surface<R>(const CtxT<R>, const StateDev, const StateDev, int);
LG_INSTANTIATE(double)
LG_INSTANTIATE(float)
#undef LG_INSTANTIATE
// the FP64 form only (the FP32 form keeps one cell per thread in these two)
template __global__ void k_diff_coef2<double>(const CtxT<double>, const StateDev);
template __global__ void k_diff_div2<double>(const CtxT<double>);

// ---- the same kernels over a group of domains (one launch per phase) ------------------------
LG_BATCH_KERNEL(diff_zeta, BX * BY, PH_DIFF_ZETA, b_diff_zeta(d.c, d.s[a.old_slot], bid))
LG_BATCH_KERNEL(diff_coef, BX * BY, PH_DIFF_COEF, b_diff_coef(d.c, d.s[a.old_slot], bid))
LG_BATCH_KERNEL(diff_dual, BX * BY, PH_DIFF_DUAL, b_diff_dual(d.c, d.s[a.old_slot], bid))
LG_BATCH_KERNEL(diff_mom, BX * BY, PH_DIFF_MOM, b_diff_mom(d.c, d.s[a.old_slot], bid))
LG_BATCH_KERNEL(diff_flux, BX * BY, PH_DIFF_FLUX, b_diff_flux(d.c, d.s[a.old_slot], bid))
LG_BATCH_KERNEL(diff_div, BX * BY, PH_DIFF_DIV, b_diff_div(d.c, bid))
LG_BATCH_KERNEL(sfc_turb, BX * BY, PH_SFC_TURB, b_sfc_turb(d.c, d.s[a.old_slot], bid))
LG_BATCH_KERNEL(pbl_wind, BX * BY, PH_PBL_WIND, b_pbl_wind(d.c, d.s[a.old_slot], bid))
LG_BATCH_KERNEL(surface, BX * BY, PH_SURFACE, b_surface(d.c, d.s[a.old_slot], d.s[a.new_slot], a.rk, bid))

template <class R>
bool batch_launch_diff(int ph, dim3 g, dim3 b, hipStream_t st, const MemberDev<R> *M, const int *prefix, int nm, const BatchArgs &a) {
#define LG_B(P, K)                                                  \
  case P:                                                           \
    hipLaunchKernelGGL(K<R>, g, b, 0, st, M, prefix, nm, a);        \
    return true;
  switch (ph) {
    LG_B(PH_DIFF_ZETA, kb_diff_zeta)
    LG_B(PH_DIFF_COEF, kb_diff_coef)
    LG_B(PH_DIFF_DUAL, kb_diff_dual)
    LG_B(PH_DIFF_MOM, kb_diff_mom)
    LG_B(PH_DIFF_FLUX, kb_diff_flux)
    LG_B(PH_DIFF_DIV, kb_diff_div)
    LG_B(PH_SFC_TURB, kb_sfc_turb)
    LG_B(PH_PBL_WIND, kb_pbl_wind)
    LG_B(PH_SURFACE, kb_surface)
    default: return false;
  }
#undef LG_B
}
template bool batch_launch_diff<double>(int, dim3, dim3, hipStream_t, const MemberDev<double> *, const int *, int, const BatchArgs &);
template bool batch_launch_diff<float>(int, dim3, dim3, hipStream_t, const MemberDev<float> *, const int *, int, const BatchArgs &);

}  // namespace lgame
