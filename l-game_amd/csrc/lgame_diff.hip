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
//   k_diff_flux_* / k_diff_div   the diffusive flux and its divergence for the temperature and
//                 for each constituent density (mo_scalar_tend_expl.F90:66-99)
// The vertical diffusion of w (mom_diff_v) is formed inside k_dissipation, which reads it.
// Every kernel keeps the reference's operation order (the library is built with
// -ffp-contract=off), so the FP64 form is bit-identical to the reference.
#include "lgame_ops.h"
#include "lgame_step.h"

namespace lgame {

// calc_diff_coeff, mo_derived.F90:189-216: molecular kinematic viscosity of air
template <class R>
__device__ __forceinline__ R calc_diff_coeff(R temperature, R density) {
  const R particle_radius = R(130e-12);
  const R particle_mass = R(M_D / N_A);
  const R thermal_velocity = sqrt(R(8.0) * R(K_B) * temperature / (R(MPI) * particle_mass));
  const R particle_density = density / particle_mass;
  const R cross_section = R(4.0) * R(MPI) * (particle_radius * particle_radius);
  const R mean_free_path = R(1.0) / (sqrt(R(2.0)) * particle_density * cross_section);
  return R(1.0 / 3.0) * thermal_velocity * mean_free_path;
}

// tke2hor_diff_coeff, mo_eff_diff_coeffs.F90:668-683
template <class R>
__device__ __forceinline__ R tke2hor(const CtxT<R> &c, R tke) {
  const R mean_velocity = sqrt(R(2.0) * tke);
  const R mean_free_path = R(c.C.eff_hor_res) / R(6.0);
  return R(1.0 / 6.0) * mean_free_path * mean_velocity;
}

// tke2vert_diff_coeff, mo_eff_diff_coeffs.F90:685-718
template <class R>
__device__ __forceinline__ R tke2vert(R tke, R n_squared, R layer_thickness) {
  const R tke_vert = R(3.0 * 1e-3) * tke;
  const R mean_velocity = sqrt(R(2.0) * tke_vert);
  const R n_used = sqrt(n_squared > R(1e-4) ? n_squared : R(1e-4));
  R mean_free_path = sqrt(R(2.0) * tke_vert) / n_used;
  mean_free_path = layer_thickness < mean_free_path ? layer_thickness : mean_free_path;
  return R(1.0 / 6.0) * mean_free_path * mean_velocity;
}

// 0.5*(sum(rho(a,1:ncond+1)) + sum(rho(b,1:ncond+1))), mo_momentum_diff_diss.F90:123-159
template <class R>
__device__ __forceinline__ R mean_rho2(const CtxT<R> &c, const StateDev &s, ix_t a, ix_t b) {
  return R(0.5) * (rho_sum(c, s, a) + rho_sum(c, s, b));
}

// ---------------------------------------------------------------------------------------------
// relative vorticity zeta_z = rel_vort_z_local/area_dual_z at all (ny+1) x (nx+1) dual points;
// row ny+1 and column nx+1 have their own metrics (mo_vorticities.F90:177-184).
// ---------------------------------------------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_zeta(const CtxT<R> c, const StateDev s) {
  KIDX_ZFAST(c.G.nx + 1, c.G.ny + 1, c.G.nl);
  if (jk > c.G.nx + 1 || ji > c.G.ny + 1 || jl > c.G.nl) return;
  AT(c.w.zeta, ji, jk, jl) = rel_vort_z(c, s, ji, jk, jl) / AT(c.g.area_dual_z, ji, jk, jl);
}

// ---------------------------------------------------------------------------------------------
// per-cell coefficients of the first substep
// ---------------------------------------------------------------------------------------------
// gravity_m_v * (d theta_v/dz) / theta_v at level L (2 <= L <= n_layers), the w_placeholder of
// update_n_squared (mo_eff_diff_coeffs.F90:636-648); th_a, th_b are theta_v of layers L-1, L.
template <class R>
__device__ __forceinline__ R n2_level(const CtxT<R> &c, int ji, int jk, int L, R th_a, R th_b) {
  const R w = (th_a - th_b) / AT(c.g.dz, ji, jk, L);
  const R inv_a = R(1.0) / th_a, inv_b = R(1.0) / th_b;
  return AT(c.g.gravity_m_v, ji, jk, L) * (R(0.5) * (inv_a + inv_b) * w);
}

template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_coef(const CtxT<R> c, const StateDev s) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (jk > nx || ji > ny || jl > nl) return;
  const Geom &G = c.G;
  const ix_t i = LG_IX(G, ji, jk, jl);
  const R rg = R(s.rho[G.ncond][i]);
  // N^2, update_n_squared (mo_eff_diff_coeffs.F90:616-666).  Its top and bottom layers take levels
  // 1 and n_levels of w_placeholder, which grad_vert does not write and which hold zero there.
  R n2 = R(0.0);
  if (jl > 1 && jl < nl) {
    const R th_m = AT(c.g.theta_v_bg, ji, jk, jl - 1) + ATR(s.theta_v_pert, ji, jk, jl - 1);
    const R th_c = c.g.theta_v_bg[i] + R(s.theta_v_pert[i]);
    const R th_p = AT(c.g.theta_v_bg, ji, jk, jl + 1) + ATR(s.theta_v_pert, ji, jk, jl + 1);
    n2 = AT(c.g.ipw[4], ji, jk, jl) * n2_level(c, ji, jk, jl, th_m, th_c) +
         AT(c.g.ipw[5], ji, jk, jl) * n2_level(c, ji, jk, jl + 1, th_c, th_p);
  }
  // tke_update (mo_tke.F90:19-67): advection by the inner product of the wind with grad(tke),
  // production by the dissipation of the previous step, decay with the viscosity of the previous
  // step, clipping at zero
  const Fld<const R> tk = c.w.tke;
  const R gx0 = grad_hor_x(c, tk, ji, jk, jl), gx1 = grad_hor_x(c, tk, ji, jk + 1, jl);
  const R gy0 = grad_hor_y(c, tk, ji, jk, jl), gy1 = grad_hor_y(c, tk, ji + 1, jk, jl);
  const VGrad<R, Fld<const R>> vg{tk};
  const R gz0 = vg(c, ji, jk, jl), gz1 = vg(c, ji, jk, jl + 1);
  const R adv = AT(c.g.ipw[0], ji, jk, jl) * ATR(s.wind_u, ji, jk + 1, jl) * gx1 + c.g.ipw[1][i] * R(s.wind_v[i]) * gy0 +
                c.g.ipw[2][i] * R(s.wind_u[i]) * gx0 + AT(c.g.ipw[3], ji, jk, jl) * ATR(s.wind_v, ji + 1, jk, jl) * gy1 +
                c.g.ipw[4][i] * R(s.wind_w[i]) * gz0 + c.g.ipw[5][i] * ATR(s.wind_w, ji, jk, jl + 1) * gz1;
  const R t_old = tk[i];
  const R decay = R(8.0 * (MPI * MPI)) / R(c.C.mva) * c.w.visc[i] / rg;
  R t = t_old + R(c.C.dtime) * (-adv + c.w.heating_diss[i] / rg - decay * t_old);
  t = t < R(0.0) ? R(0.0) : t;
  // hor_viscosity (mo_eff_diff_coeffs.F90:24-159): molecular plus eddy viscosity, times rho
  const R mol = calc_diff_coeff(c.w.temperature[i], rg);
  const R div = div_h(c, Fld<const double>(s.wind_u), Fld<const double>(s.wind_v), ji, jk, jl);
  R eddy;
  if (!c.C.scheme_tke) {
    const Fld<const R> z = c.w.zeta;
    const R vc = R(0.25) * (AT(z, ji, jk, jl) + AT(z, ji, jk + 1, jl) + AT(z, ji + 1, jk, jl) + AT(z, ji + 1, jk + 1, jl));
    const R a = vc * vc + div * div, b = R(c.C.bg_shear) * R(c.C.bg_shear);
    eddy = R(c.C.c_s) * R(c.C.mva) * sqrt(a > b ? a : b);
  } else {
    eddy = tke2hor(c, t);
  }
  const R visc = rg * (mol + eddy);
  c.w.tke_next[i] = t;
  c.w.visc[i] = visc;
  c.w.kv[i] = mol + tke2vert(t, n2, c.g.layer_thickness[i]);
  c.w.vdiv[i] = visc * div;
}

// ---------------------------------------------------------------------------------------------
// viscosity_dual * zeta_z at the dual points.  viscosity_dual averages the four cells around the
// dual point and is multiplied by their average gas density; each edge and corner keeps the
// reference's order of the four terms (mo_eff_diff_coeffs.F90:214-307).  viscosity_dual of row
// ny+1 and column nx+1 is that of row 1 and column 1; zeta_z there is their own.
// ---------------------------------------------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_dual(const CtxT<R> c, const StateDev s) {
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

// ---------------------------------------------------------------------------------------------
// mom_diff_h (mo_momentum_diff_diss.F90:25-160): the gradient of viscosity times divergence plus
// the curl of viscosity_dual times vorticity with its terrain-following correction, divided by
// the averaged density, at the u and v points of each cell.  The terrain correction of the
// gradient reads mom_diff_tend_z as the previous step left it (grad_hor at :62 takes it as the
// vertical component).  Column nx+1 and row ny+1 are copies; the TKE of this step replaces the
// carried one.
// ---------------------------------------------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_mom(const CtxT<R> c, const StateDev s) {
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
    const R mx = (gx + u) / mean_rho2(c, s, LG_IX(G, ji, km, jl), i);
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
    const R my = (gy + v) / mean_rho2(c, s, LG_IX(G, jm, jk, jl), i);
    c.w.mom_y[i] = my;
    if (ji == 1) AT(c.w.mom_y, ny + 1, jk, jl) = my;
  }
  c.w.tke[i] = c.w.tke_next[i];
}

// ---------------------------------------------------------------------------------------------
// scalar diffusion (mo_scalar_tend_expl.F90:66-99 with scalar_diff_coeffs,
// mo_eff_diff_coeffs.F90:526-614): the flux coefficient times grad_hor / grad_vert of the field
// at the u, v and w points of each cell, then the divergence.  Coefficients: temperature
// c_g_v*viscosity and rho_gas*c_g_v*kv, constituent densities viscosity/rho_gas and kv; c_g_v is
// spec_heat_cap_diagnostics_v (mo_derived.F90:40-68) with its one relevant constituent.
// ---------------------------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ R c_g_v(R rg) {
  R r = R(0.0);
  r = r + rg / rg * R(C_D_V);
  return r;
}

template <class R, bool TEMP, class P>
__device__ __forceinline__ void diff_flux(const CtxT<R> &c, const StateDev &s, P f, bool lvert, int ji, int jk, int jl) {
  const Geom &G = c.G;
  const int nx = G.nx, ny = G.ny;
  const Fld<const double> rgf = s.rho[G.ncond];
  auto ch = [&](ix_t q) {
    const R rg = R(rgf[q]);
    return TEMP ? c_g_v(rg) * c.w.visc[q] : c.w.visc[q] / rg;
  };
  auto cv = [&](ix_t q) {
    const R rg = R(rgf[q]);
    return TEMP ? rg * c_g_v(rg) * c.w.kv[q] : c.w.kv[q];
  };
  const ix_t i = LG_IX(G, ji, jk, jl);
  const R ch_c = ch(i);
  {
    const int km = wrapm(jk, nx);
    const R fu = R(0.5) * (ch(LG_IX(G, ji, km, jl)) + ch_c) * grad_hor_x(c, f, ji, jk, jl);
    c.w.dfu[i] = fu;
    if (jk == 1) AT(c.w.dfu, ji, nx + 1, jl) = fu;
  }
  {
    const int jm = wrapm(ji, ny);
    const R fv = R(0.5) * (ch(LG_IX(G, jm, jk, jl)) + ch_c) * grad_hor_y(c, f, ji, jk, jl);
    c.w.dfv[i] = fv;
    if (ji == 1) AT(c.w.dfv, ny + 1, jk, jl) = fv;
  }
  if (lvert && jl >= 2) {
    const VGrad<R, P> vg{f};
    c.w.dfw[i] = R(0.5) * (cv(LG_IX(G, ji, jk, jl - 1)) + cv(i)) * vg(c, ji, jk, jl);
  }
}

template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_flux_temp(const CtxT<R> c, const StateDev s) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  diff_flux<R, true>(c, s, Fld<const R>(c.w.temperature), c.C.ltemp_v != 0, ji, jk, jl);
}

template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_flux_mass(const CtxT<R> c, const StateDev s, int jc) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  diff_flux<R, false>(c, s, Fld<const double>(s.rho[jc]), c.C.lmass_v != 0, ji, jk, jl);
}

// div_h of the horizontal fluxes plus add_vertical_div of the vertical one
// (mo_divergence_operators.F90:16-68,149-186)
template <class R>
__global__ void __launch_bounds__(BX * BY) k_diff_div(const CtxT<R> c, Fld<R> out, int lvert) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  const int nl = c.G.nl;
  if (jk > c.G.nx || ji > c.G.ny || jl > nl) return;
  const ix_t i = LG_IX(c.G, ji, jk, jl);
  R d = div_h(c, Fld<const R>(c.w.dfu), Fld<const R>(c.w.dfv), ji, jk, jl);
  if (lvert) {
    const R cu = jl == 1 ? R(0.0) : c.w.dfw[i];
    const R cl = jl == nl ? R(0.0) : AT(c.w.dfw, ji, jk, jl + 1);
    const R comp_v = cu * c.g.area_z[i] - cl * AT(c.g.area_z, ji, jk, jl + 1);
    d = d + comp_v / c.g.volume[i];
  }
  out[i] = d;
}

#define LG_INSTANTIATE(R)                                                                  \
  template __global__ void k_diff_zeta<R>(const CtxT<R>, const StateDev);                  \
  template __global__ void k_diff_coef<R>(const CtxT<R>, const StateDev);                  \
  template __global__ void k_diff_dual<R>(const CtxT<R>, const StateDev);                  \
  template __global__ void k_diff_mom<R>(const CtxT<R>, const StateDev);                   \
  template __global__ void k_diff_flux_temp<R>(const CtxT<R>, const StateDev);             \
  template __global__ void k_diff_flux_mass<R>(const CtxT<R>, const StateDev, int);        \
  template __global__ void k_diff_div<R>(const CtxT<R>, Fld<R>, int);
LG_INSTANTIATE(double)
LG_INSTANTIATE(float)
#undef LG_INSTANTIATE

}  // namespace lgame
