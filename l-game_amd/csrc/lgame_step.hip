// lgame_step.hip -- HIP kernels of the HEVI predictor-corrector step for gfx950 (FP64 form).
//
// One kernel per fused operator group of pchevi (src/time_stepping/mo_manage_pchevi.F90:27-118).
// Horizontal stencil kernels map threadIdx.x to x (64 consecutive x points per wavefront, the
// contiguous direction of the layout), threadIdx.y to y and blockIdx.z to the level.  The two
// implicit column solvers map one column per lane and keep the Thomas recursion in registers
// with per-level values in coalesced column scratch.
#include "lgame_ops.h"
#include "lgame_step.h"

namespace lgame {

#define KIDX                                                        \
  const int jk = blockIdx.x * blockDim.x + threadIdx.x + 1;         \
  const int ji = blockIdx.y * blockDim.y + threadIdx.y + 1;         \
  const int jl = blockIdx.z + 1;

// ---------------------------------------------------------------------------------------------
// temperature_diagnostics, mo_derived.F90:18-38
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(BX * BY) k_temperature(const Ctx c, const StateDev s) {
  KIDX;
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  const size_t i = LG_IX(c.G, ji, jk, jl);
  double t = (c.g.theta_v_bg[i] + s.theta_v_pert[i]) * (c.g.exner_bg[i] + s.exner_pert[i]);
  if (c.C.lmoist) t = t / (1.0 + s.rho[c.G.ncond + 1][i] / s.rho[c.G.ncond][i] * (M_D / M_V - 1.0));
  c.w.temperature[i] = t;
}

// ---------------------------------------------------------------------------------------------
// theta_v_adv_3rd_order, mo_multiplications.F90:116-204, in three passes.
// field f: 0 = full theta_v (theta_v_bg + theta_v_pert), 1 = theta_v_pert.
// ---------------------------------------------------------------------------------------------
// pass 1a: scalar_placeholder = theta_v_bg + theta_v_pert (f=0) or theta_v_pert (f=1)
__global__ void __launch_bounds__(BX * BY) k_theta_fill(const Ctx c, const StateDev s, int nfield) {
  KIDX;
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  const size_t i = LG_IX(c.G, ji, jk, jl);
  c.w.tth[0][i] = c.g.theta_v_bg[i] + s.theta_v_pert[i];
  if (nfield > 1) c.w.tth[1][i] = s.theta_v_pert[i];
}

// pass 1b: u,v_placeholder = grad_hor(theta) with its grad_vert inline
__global__ void __launch_bounds__(BX * BY) k_theta_grad(const Ctx c, int nfield) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny;
  if (jk > nx + 1 || ji > ny + 1 || jl > c.G.nl) return;
  for (int f = 0; f < nfield; ++f) {
    if (ji <= ny) AT(c.w.tgu[f], ji, jk, jl) = grad_hor_x(c, c.w.tth[f], ji, jk, jl);
    if (jk <= nx) AT(c.w.tgv[f], ji, jk, jl) = grad_hor_y(c, c.w.tth[f], ji, jk, jl);
  }
}

// pass 2: flux_density_div = div_h(u_placeholder, v_placeholder)
__global__ void __launch_bounds__(BX * BY) k_theta_div(const Ctx c, int nfield) {
  KIDX;
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  for (int f = 0; f < nfield; ++f) AT(c.w.tfd[f], ji, jk, jl) = div_h(c, c.w.tgu[f], c.w.tgv[f], ji, jk, jl);
}

// pass 3: upwind-corrected edge values; the upwind direction comes from the given winds
__global__ void __launch_bounds__(BX * BY) k_theta_edge(const Ctx c, const double *wind_u, const double *wind_v, int nfield) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny;
  if (jk > nx + 1 || ji > ny + 1 || jl > c.G.nl) return;
  const double c8 = 1.0 / 8.0;
  for (int f = 0; f < nfield; ++f) {
    const double *th = c.w.tth[f];
    const double *fd = c.w.tfd[f];
    if (ji <= ny) {
      const int k = (jk == nx + 1) ? 1 : jk;  // theta_v_u(:,nx+1,:) = theta_v_u(:,1,:)
      const int km = wrapm(k, nx);
      double t = 0.5 * (AT(th, ji, km, jl) + AT(th, ji, k, jl));
      if (AT(wind_u, ji, k, jl) >= 0.0)
        t = t - c8 * c.C.dyn * c.C.dxn * AT(fd, ji, km, jl);
      else
        t = t - c8 * c.C.dyn * c.C.dxn * AT(fd, ji, k, jl);
      AT(c.w.tu[f], ji, jk, jl) = t;
    }
    if (jk <= nx) {
      const int j = (ji == ny + 1) ? 1 : ji;
      const int jm = wrapm(j, ny);
      double t = 0.5 * (AT(th, j, jk, jl) + AT(th, jm, jk, jl));
      if (AT(wind_v, j, jk, jl) >= 0.0)
        t = t - c8 * c.C.dyn * c.C.dxn * AT(fd, j, jk, jl);
      else
        t = t - c8 * c.C.dyn * c.C.dxn * AT(fd, jm, jk, jl);
      AT(c.w.tv[f], ji, jk, jl) = t;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// manage_p_grad, mo_p_grad.F90:20-104 (luse_bg_state): nonlinear and linear pressure gradient
// accelerations at u, v and w points, scaled by the condensate deceleration factor.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double decel(const Ctx &c, const StateDev &s, int ji, int jk, int jl) {
  const size_t i = LG_IX(c.G, ji, jk, jl);
  return C_D_P * s.rho[c.G.ncond][i] / rho_sum(c, s, i);
}

__global__ void __launch_bounds__(BX * BY) k_pgrad(const Ctx c, const StateDev s, int lfirst) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (jk > nx + 1 || ji > ny + 1 || jl > nl + 1) return;
  const bool o3 = c.C.order == 3;
  if (ji <= ny && jl <= nl) {  // u point; the final value at nx+1 is the copy of jk = 1
    const size_t iu = LG_IX(c.G, ji, jk, jl);
    if (!lfirst) c.w.pold_u[iu] = -c.w.pnl_u[iu] - 1.0 * c.w.pl_u[iu];
    const int k = (jk == nx + 1) ? 1 : jk, km = wrapm(k, nx);
    const double g = grad_hor_x(c, s.exner_pert, ji, k, jl);
    double t, lu;
    if (o3) {
      t = AT(c.w.tu[0], ji, k, jl) * g;
      lu = AT(c.w.tu[1], ji, k, jl) * AT(c.g.ebg_u, ji, k, jl);
    } else {
      t = 0.5 * ((AT(c.g.theta_v_bg, ji, km, jl) + AT(s.theta_v_pert, ji, km, jl)) +
                 (AT(c.g.theta_v_bg, ji, k, jl) + AT(s.theta_v_pert, ji, k, jl))) * g;
      lu = 0.5 * (AT(s.theta_v_pert, ji, km, jl) + AT(s.theta_v_pert, ji, k, jl)) * AT(c.g.ebg_u, ji, k, jl);
    }
    const double d = 0.5 * (decel(c, s, ji, km, jl) + decel(c, s, ji, k, jl));
    const double nlu = d * t, llu = d * lu;
    c.w.pnl_u[iu] = nlu;
    c.w.pl_u[iu] = llu;
    if (lfirst) c.w.pold_u[iu] = -nlu - 1.0 * llu;
  }
  if (jk <= nx && jl <= nl) {  // v point
    const size_t iv = LG_IX(c.G, ji, jk, jl);
    if (!lfirst) c.w.pold_v[iv] = -c.w.pnl_v[iv] - 1.0 * c.w.pl_v[iv];
    const int j = (ji == ny + 1) ? 1 : ji, jm = wrapm(j, ny);
    const double g = grad_hor_y(c, s.exner_pert, j, jk, jl);
    double t, lv;
    if (o3) {
      t = AT(c.w.tv[0], j, jk, jl) * g;
      lv = AT(c.w.tv[1], j, jk, jl) * AT(c.g.ebg_v, j, jk, jl);
    } else {
      t = 0.5 * ((AT(c.g.theta_v_bg, jm, jk, jl) + AT(s.theta_v_pert, jm, jk, jl)) +
                 (AT(c.g.theta_v_bg, j, jk, jl) + AT(s.theta_v_pert, j, jk, jl))) * g;
      lv = 0.5 * (AT(s.theta_v_pert, jm, jk, jl) + AT(s.theta_v_pert, j, jk, jl)) * AT(c.g.ebg_v, j, jk, jl);
    }
    const double d = 0.5 * (decel(c, s, jm, jk, jl) + decel(c, s, j, jk, jl));
    const double nlv = d * t, llv = d * lv;
    c.w.pnl_v[iv] = nlv;
    c.w.pl_v[iv] = llv;
    if (lfirst) c.w.pold_v[iv] = -nlv - 1.0 * llv;
  }
  if (ji <= ny && jk <= nx && jl >= 2 && jl <= nl) {  // w point, levels 2..n_layers
    const size_t iw = LG_IX(c.G, ji, jk, jl);
    const size_t ia = LG_IX(c.G, ji, jk, jl - 1);
    const double g = (s.exner_pert[ia] - s.exner_pert[iw]) / c.g.dz[iw];
    const double t = 0.5 * ((c.g.theta_v_bg[ia] + s.theta_v_pert[ia]) + (c.g.theta_v_bg[iw] + s.theta_v_pert[iw])) * g;
    const double lw = 0.5 * (s.theta_v_pert[ia] + s.theta_v_pert[iw]) * c.g.ebg_w[iw];
    const double d = 0.5 * (decel(c, s, ji, jk, jl - 1) + decel(c, s, ji, jk, jl));
    c.w.pnl_w[iw] = d * t;
    c.w.pl_w[iw] = d * lw;
  }
}

// ---------------------------------------------------------------------------------------------
// Momentum advection block of vector_tend_expl (mo_vector_tend_expl.F90:58-99):
// potential vorticity, mass fluxes, vorticity flux term, kinetic-energy gradient.
// ---------------------------------------------------------------------------------------------

// circulation of zeta_x around its dual face, mo_vorticities.F90:33-93 (jl in 2..n_levels)
__device__ double zeta_x_circ(const Ctx &c, const StateDev &s, int ji, int jk, int jl) {
  const int ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev;
  const int jm = wrapm(ji, ny);
  if (jl <= nl) {
    return AT(c.g.dz, jm, jk, jl) * AT(s.wind_w, jm, jk, jl) - AT(c.g.dy, ji, jk, jl - 1) * hcov_y(c, s.wind_v, s.wind_w, ji, jk, jl - 1) -
           AT(c.g.dz, ji, jk, jl) * AT(s.wind_w, ji, jk, jl) + AT(c.g.dy, ji, jk, jl) * hcov_y(c, s.wind_v, s.wind_w, ji, jk, jl);
  }
  double r = -AT(c.g.dy, ji, jk, nl) * hcov_y(c, s.wind_v, s.wind_w, ji, jk, nl);
  if (c.C.lfreeslip) {
    r = r + 0.5 * AT(c.g.dz, jm, jk, nlev) * AT(s.wind_w, jm, jk, nlev) - 0.5 * AT(c.g.dz, ji, jk, nlev) * AT(s.wind_w, ji, jk, nlev) +
        AT(c.g.dy, ji, jk, nl) *
            (AT(s.wind_v, ji, jk, nl) +
             (AT(s.wind_v, ji, jk, nl - 1) - AT(s.wind_v, ji, jk, nl)) / (AT(c.g.z_v, ji, jk, nl - 1) - AT(c.g.z_v, ji, jk, nl)) *
                 (0.5 * (AT(c.g.z_w, jm, jk, nlev) + AT(c.g.z_w, ji, jk, nlev)) - AT(c.g.z_v, ji, jk, nl)) +
             0.5 * (AT(s.wind_w, jm, jk, nlev) + AT(s.wind_w, ji, jk, nlev)) * (AT(c.g.z_w, jm, jk, nlev) - AT(c.g.z_w, ji, jk, nlev)) /
                 AT(c.g.dy, ji, jk, nl));
  }
  return r;
}

// mo_vorticities.F90:105-165 (jl in 2..n_levels)
__device__ double zeta_y_circ(const Ctx &c, const StateDev &s, int ji, int jk, int jl) {
  const int nx = c.G.nx, nl = c.G.nl, nlev = c.G.nlev;
  const int km = wrapm(jk, nx);
  if (jl <= nl) {
    return -AT(c.g.dz, ji, jk, jl) * AT(s.wind_w, ji, jk, jl) - AT(c.g.dx, ji, jk, jl) * hcov_x(c, s.wind_u, s.wind_w, ji, jk, jl) +
           AT(c.g.dz, ji, km, jl) * AT(s.wind_w, ji, km, jl) + AT(c.g.dx, ji, jk, jl - 1) * hcov_x(c, s.wind_u, s.wind_w, ji, jk, jl - 1);
  }
  double r = AT(c.g.dx, ji, jk, nl) * hcov_x(c, s.wind_u, s.wind_w, ji, jk, nl);
  if (c.C.lfreeslip) {
    r = r + 0.5 * AT(c.g.dz, ji, km, nlev) * AT(s.wind_w, ji, km, nlev) - 0.5 * AT(c.g.dz, ji, jk, nlev) * AT(s.wind_w, ji, jk, nlev) -
        AT(c.g.dx, ji, jk, nl) *
            (AT(s.wind_u, ji, jk, nl) +
             (AT(s.wind_u, ji, jk, nl - 1) - AT(s.wind_u, ji, jk, nl)) / (AT(c.g.z_u, ji, jk, nl - 1) - AT(c.g.z_u, ji, jk, nl)) *
                 (0.5 * (AT(c.g.z_w, ji, km, nlev) + AT(c.g.z_w, ji, jk, nlev)) - AT(c.g.z_u, ji, jk, nl)) +
             0.5 * (AT(s.wind_w, ji, km, nlev) + AT(s.wind_w, ji, jk, nlev)) * (AT(c.g.z_w, ji, jk, nlev) - AT(c.g.z_w, ji, km, nlev)) /
                 AT(c.g.dx, ji, jk, nl));
  }
  return r;
}

// rel_vort_z_local, mo_vorticities.F90:423-538, for a canonical dual point (ji<=ny, jk<=nx)
__device__ double rel_vort_z(const Ctx &c, const StateDev &s, int ji, int jk, int jl) {
  const int ny = c.G.ny, nx = c.G.nx;
  const int j_i[4] = {ji, wrapm(ji, ny), ji, ji};
  const int j_k[4] = {jk, jk, wrapm(jk, nx), jk};
  const double sgn[4] = {1.0, -1.0, -1.0, 1.0};
  double r = 0.0;
  if (jl <= c.G.nflat) {
    r = r + sgn[0] * AT(c.g.dy, j_i[0], j_k[0], jl) * AT(s.wind_v, j_i[0], j_k[0], jl) +
        sgn[1] * AT(c.g.dx, j_i[1], j_k[1], jl) * AT(s.wind_u, j_i[1], j_k[1], jl) +
        sgn[2] * AT(c.g.dy, j_i[2], j_k[2], jl) * AT(s.wind_v, j_i[2], j_k[2], jl) +
        sgn[3] * AT(c.g.dx, j_i[3], j_k[3], jl) * AT(s.wind_u, j_i[3], j_k[3], jl);
    return r;
  }
  const double zd = AT(c.g.z_area_dual_z, ji, jk, jl);
#pragma unroll
  for (int jm = 0; jm < 4; ++jm) {
    const bool isv = (jm == 0 || jm == 2);
    const double *zf = isv ? c.g.z_v : c.g.z_u;
    const double *wf = isv ? s.wind_v : s.wind_u;
    const double *len = isv ? c.g.dy : c.g.dx;
    const double zp = AT(zf, j_i[jm], j_k[jm], jl);
    const double l_rescale = (R_E + zd) / (R_E + zp);
    const double delta_z = zd - zp;
    int sh = 1;
    if (delta_z > 0.0 || jl == c.G.nl) sh = -1;
    if (jl == 1) sh = 1;
    const double vg = (AT(wf, j_i[jm], j_k[jm], jl) - AT(wf, j_i[jm], j_k[jm], jl + sh)) / (zp - AT(zf, j_i[jm], j_k[jm], jl + sh));
    r = r + sgn[jm] * l_rescale * AT(len, j_i[jm], j_k[jm], jl) * (AT(wf, j_i[jm], j_k[jm], jl) + delta_z * vg);
  }
  return r;
}

__device__ __forceinline__ double rhog(const Ctx &c, const StateDev &s, int ji, int jk, int jl) {
  return AT(s.rho[c.G.ncond], ji, jk, jl);
}

// density average of calc_pot_vort for horizontal vorticities, mo_vorticities.F90:240-266
__device__ double pv_den_h(const Ctx &c, const StateDev &s, int ia, int ka, int ib, int kb, int jl) {
  if (jl == 1) {
    return 0.5 * (rhog(c, s, ia, ka, jl) +
                  (c.C.toa - AT(c.g.z_scalar, ia, ka, jl)) * (rhog(c, s, ia, ka, jl) - rhog(c, s, ia, ka, jl + 1)) /
                      (AT(c.g.z_scalar, ia, ka, jl) - AT(c.g.z_scalar, ia, ka, jl + 1)) +
                  rhog(c, s, ib, kb, jl) +
                  (c.C.toa - AT(c.g.z_scalar, ib, kb, jl)) * (rhog(c, s, ib, kb, jl) - rhog(c, s, ib, kb, jl + 1)) /
                      (AT(c.g.z_scalar, ib, kb, jl) - AT(c.g.z_scalar, ib, kb, jl + 1)));
  } else if (jl == c.G.nlev) {
    return 0.5 * (rhog(c, s, ia, ka, jl - 1) +
                  (AT(c.g.z_w, ia, ka, jl) - AT(c.g.z_scalar, ia, ka, jl - 1)) * (rhog(c, s, ia, ka, jl - 2) - rhog(c, s, ia, ka, jl - 1)) /
                      (AT(c.g.z_scalar, ia, ka, jl - 2) - AT(c.g.z_scalar, ia, ka, jl - 1)) +
                  rhog(c, s, ib, kb, jl - 1) +
                  (AT(c.g.z_w, ib, kb, jl) - AT(c.g.z_scalar, ib, kb, jl - 1)) * (rhog(c, s, ib, kb, jl - 2) - rhog(c, s, ib, kb, jl - 1)) /
                      (AT(c.g.z_scalar, ib, kb, jl - 2) - AT(c.g.z_scalar, ib, kb, jl - 1)));
  }
  return 0.25 * (rhog(c, s, ia, ka, jl - 1) + rhog(c, s, ib, kb, jl - 1) + rhog(c, s, ia, ka, jl) + rhog(c, s, ib, kb, jl));
}

// calc_pot_vort (mo_vorticities.F90:189-421) with rel_vort (:19-187) inline
__global__ void __launch_bounds__(BX * BY) k_pot_vort(const Ctx c, const StateDev s) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev;
  if (jk > nx + 1 || ji > ny + 1 || jl > nlev) return;
  if (jk <= nx) {  // eta_x; row ny+1 is the copy of row 1
    const int j = (ji == ny + 1) ? 1 : ji;
    double z = 0.0;
    if (!c.C.llinear) {
      const int le = jl == 1 ? 2 : jl;  // zeta_x(:,:,1) = zeta_x(:,:,2)
      z = zeta_x_circ(c, s, j, jk, le) / AT(c.g.area_dual_x, j, jk, le);
    }
    double e = c.C.lcorio ? z + c.g.fvec_x[LG_IX(c.G, j, jk, 1)] : z;
    AT(c.w.eta_x, ji, jk, jl) = e / pv_den_h(c, s, wrapm(j, ny), jk, j, jk, jl);
  }
  if (ji <= ny) {  // eta_y; column nx+1 is the copy of column 1
    const int k = (jk == nx + 1) ? 1 : jk;
    double z = 0.0;
    if (!c.C.llinear) {
      const int le = jl == 1 ? 2 : jl;
      z = zeta_y_circ(c, s, ji, k, le) / AT(c.g.area_dual_y, ji, k, le);
    }
    double e = c.C.lcorio ? z + c.g.fvec_y[LG_IX(c.G, ji, k, 1)] : z;
    AT(c.w.eta_y, ji, jk, jl) = e / pv_den_h(c, s, ji, wrapm(k, nx), ji, k, jl);
  }
  if (jl <= nl) {  // eta_z; duplicates are copies of the canonical point
    const int j = (ji == ny + 1) ? 1 : ji, k = (jk == nx + 1) ? 1 : jk;
    const int jm = wrapm(j, ny), km = wrapm(k, nx);
    double z = 0.0;
    if (!c.C.llinear) z = rel_vort_z(c, s, j, k, jl) / AT(c.g.area_dual_z, j, k, jl);
    double e = c.C.lcorio ? z + c.g.fvec_z[LG_IX(c.G, j, k, 1)] : z;
    AT(c.w.eta_z, ji, jk, jl) =
        e / (0.25 * (rhog(c, s, j, km, jl) + rhog(c, s, j, k, jl) + rhog(c, s, jm, k, jl) + rhog(c, s, jm, km, jl)));
  }
}

// mass flux densities of the gas, scalar_times_vector_h/v (mo_vector_tend_expl.F90:64-80)
__global__ void __launch_bounds__(BX * BY) k_mass_flux(const Ctx c, const StateDev s) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev;
  if (jk > nx + 1 || ji > ny + 1 || jl > nlev) return;
  if (ji <= ny && jl <= nl) {
    const int k = (jk == nx + 1) ? 1 : jk, km = wrapm(k, nx);
    AT(c.w.u_ph, ji, jk, jl) = 0.5 * (rhog(c, s, ji, km, jl) + rhog(c, s, ji, k, jl)) * AT(s.wind_u, ji, k, jl);
  }
  if (jk <= nx && jl <= nl) {
    const int j = (ji == ny + 1) ? 1 : ji, jm = wrapm(j, ny);
    AT(c.w.v_ph, ji, jk, jl) = 0.5 * (rhog(c, s, jm, jk, jl) + rhog(c, s, j, jk, jl)) * AT(s.wind_v, j, jk, jl);
  }
  if (ji <= ny && jk <= nx) {
    double r = 0.0;
    if (jl >= 2 && jl <= nl) {
      r = 0.5 * (rhog(c, s, ji, jk, jl - 1) + rhog(c, s, ji, jk, jl)) * AT(s.wind_w, ji, jk, jl);
    } else if (jl == nlev && c.C.lfreeslip) {
      const double dv = rhog(c, s, ji, jk, nl) + (rhog(c, s, ji, jk, nl - 1) - rhog(c, s, ji, jk, nl)) /
                                                     (AT(c.g.z_scalar, ji, jk, nl - 1) - AT(c.g.z_scalar, ji, jk, nl)) *
                                                     (AT(c.g.z_w, ji, jk, nlev) - AT(c.g.z_scalar, ji, jk, nl));
      r = dv * AT(s.wind_w, ji, jk, nlev);
    }
    AT(c.w.w_ph, ji, jk, jl) = r;
  }
}

// calc_vorticity_flux_term, mo_vorticity_flux.F90:16-171
__global__ void __launch_bounds__(BX * BY) k_vort_flux(const Ctx c) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (jk > nx + 1 || ji > ny + 1 || jl > nl) return;
  const double *ez = c.w.eta_z, *up = c.w.u_ph, *vp = c.w.v_ph, *wp = c.w.w_ph;
  if (ji <= ny) {  // u points; nx+1 is the copy of column 1
    const int k = (jk == nx + 1) ? 1 : jk, km = wrapm(k, nx), kp = k + 1;
    const double *tw = c.g.trsk_u + 3 * (ji - 1);
    double r = tw[0] * AT(vp, ji, km, jl) * 0.25 * (AT(ez, ji, km, jl) + AT(ez, ji, k, jl) + AT(ez, ji, k, jl) + AT(ez, ji + 1, k, jl)) +
               tw[1] * AT(up, ji, km, jl) * 0.25 * (AT(ez, ji, km, jl) + AT(ez, ji + 1, km, jl) + AT(ez, ji, k, jl) + AT(ez, ji + 1, k, jl)) +
               tw[2] * AT(vp, ji + 1, km, jl) * 0.25 * (AT(ez, ji + 1, km, jl) + AT(ez, ji + 1, k, jl) + AT(ez, ji, k, jl) + AT(ez, ji + 1, k, jl)) +
               tw[2] * AT(vp, ji + 1, k, jl) * 0.25 * (AT(ez, ji + 1, k, jl) + AT(ez, ji + 1, kp, jl) + AT(ez, ji, k, jl) + AT(ez, ji + 1, k, jl)) +
               tw[1] * AT(up, ji, kp, jl) * 0.25 * (AT(ez, ji + 1, kp, jl) + AT(ez, ji, kp, jl) + AT(ez, ji, k, jl) + AT(ez, ji + 1, k, jl)) +
               tw[0] * AT(vp, ji, k, jl) * 0.25 * (AT(ez, ji, k, jl) + AT(ez, ji, kp, jl) + AT(ez, ji, k, jl) + AT(ez, ji + 1, k, jl));
    r = r - 0.5 * AT(c.g.ipw[4], ji, km, jl) * AT(wp, ji, km, jl) * AT(c.w.eta_y, ji, k, jl) -
        0.5 * AT(c.g.ipw[5], ji, km, jl) * AT(wp, ji, km, jl + 1) * AT(c.w.eta_y, ji, k, jl + 1) -
        0.5 * AT(c.g.ipw[4], ji, k, jl) * AT(wp, ji, k, jl) * AT(c.w.eta_y, ji, k, jl) -
        0.5 * AT(c.g.ipw[5], ji, k, jl) * AT(wp, ji, k, jl + 1) * AT(c.w.eta_y, ji, k, jl + 1);
    AT(c.w.pvt_x, ji, jk, jl) = r;
  }
  if (jk <= nx) {  // v points; ny+1 is the copy of row 1
    const int j = (ji == ny + 1) ? 1 : ji, jm = wrapm(j, ny);
    const double *tw = c.g.trsk_v + 2 * (j - 1);
    double r = tw[0] * AT(up, j, jk, jl) * 0.25 * (AT(ez, j, jk, jl) + AT(ez, j + 1, jk, jl) + AT(ez, j, jk, jl) + AT(ez, j, jk + 1, jl)) +
               tw[0] * AT(up, j, jk + 1, jl) * 0.25 * (AT(ez, j, jk + 1, jl) + AT(ez, j + 1, jk + 1, jl) + AT(ez, j, jk, jl) + AT(ez, j, jk + 1, jl)) +
               tw[1] * AT(up, jm, jk + 1, jl) * 0.25 * (AT(ez, jm, jk + 1, jl) + AT(ez, j, jk + 1, jl) + AT(ez, j, jk, jl) + AT(ez, j, jk + 1, jl)) +
               tw[1] * AT(up, jm, jk, jl) * 0.25 * (AT(ez, jm, jk, jl) + AT(ez, j, jk, jl) + AT(ez, j, jk, jl) + AT(ez, j, jk + 1, jl));
    r = r + 0.5 * AT(c.g.ipw[4], jm, jk, jl) * AT(wp, jm, jk, jl) * AT(c.w.eta_x, j, jk, jl) +
        0.5 * AT(c.g.ipw[5], jm, jk, jl) * AT(wp, jm, jk, jl + 1) * AT(c.w.eta_x, j, jk, jl + 1) +
        0.5 * AT(c.g.ipw[4], j, jk, jl) * AT(wp, j, jk, jl) * AT(c.w.eta_x, j, jk, jl) +
        0.5 * AT(c.g.ipw[5], j, jk, jl) * AT(wp, j, jk, jl + 1) * AT(c.w.eta_x, j, jk, jl + 1);
    AT(c.w.pvt_y, ji, jk, jl) = r;
  }
  if (ji <= ny && jk <= nx && jl >= 2) {  // w points, levels 2..n_layers
    const int l = jl - 1;
    AT(c.w.pvt_z, ji, jk, jl) =
        0.5 * (AT(c.g.ipw[0], ji, jk, l) * AT(up, ji, jk + 1, l) * AT(c.w.eta_y, ji, jk + 1, jl) -
               AT(c.g.ipw[1], ji, jk, l) * AT(vp, ji, jk, l) * AT(c.w.eta_x, ji, jk, jl) +
               AT(c.g.ipw[2], ji, jk, l) * AT(up, ji, jk, l) * AT(c.w.eta_y, ji, jk, jl) -
               AT(c.g.ipw[3], ji, jk, l) * AT(vp, ji + 1, jk, l) * AT(c.w.eta_x, ji + 1, jk, jl) +
               AT(c.g.ipw[0], ji, jk, jl) * AT(up, ji, jk + 1, jl) * AT(c.w.eta_y, ji, jk + 1, jl) -
               AT(c.g.ipw[1], ji, jk, jl) * AT(vp, ji, jk, jl) * AT(c.w.eta_x, ji, jk, jl) +
               AT(c.g.ipw[2], ji, jk, jl) * AT(up, ji, jk, jl) * AT(c.w.eta_y, ji, jk, jl) -
               AT(c.g.ipw[3], ji, jk, jl) * AT(vp, ji + 1, jk, jl) * AT(c.w.eta_x, ji + 1, jk, jl));
  }
}

__global__ void __launch_bounds__(BX * BY) k_ke(const Ctx c, const StateDev s) {
  KIDX;
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  AT(c.w.v_squared, ji, jk, jl) = ke_point(c, s, ji, jk, jl);
}

// grad_vert + grad_hor of the kinetic energy (mo_vector_tend_expl.F90:96-97)
__global__ void __launch_bounds__(BX * BY) k_ke_grad(const Ctx c) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (jk > nx + 1 || ji > ny + 1 || jl > nl) return;
  const double *k2 = c.w.v_squared;
  if (ji <= ny) AT(c.w.vsg_x, ji, jk, jl) = grad_hor_x(c, k2, ji, jk, jl);
  if (jk <= nx) AT(c.w.vsg_y, ji, jk, jl) = grad_hor_y(c, k2, ji, jk, jl);
  if (ji <= ny && jk <= nx && jl >= 2)
    AT(c.w.vsg_z, ji, jk, jl) = (AT(k2, ji, jk, jl - 1) - AT(k2, ji, jk, jl)) / AT(c.g.dz, ji, jk, jl);
}

// simple_dissipation_rate, mo_momentum_diff_diss.F90:333-377 (momentum diffusion off: only the
// Klemp layer acceleration enters the inner product)
__global__ void __launch_bounds__(BX * BY) k_dissipation(const Ctx c, const StateDev s) {
  KIDX;
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  const size_t i = LG_IX(c.G, ji, jk, jl), ib = LG_IX(c.G, ji, jk, jl + 1);
  double wk0 = 0.0, wk1 = 0.0;
  const double dt = c.C.dtime;
  if (jl >= 2 && jl <= c.C.n_damping_levels) {
    const double w = s.wind_w[i];
    wk0 = (w / c.g.klemp_pref[i] - w) / dt;
  }
  if (jl + 1 >= 2 && jl + 1 <= c.C.n_damping_levels) {
    const double w = s.wind_w[ib];
    wk1 = (w / c.g.klemp_pref[ib] - w) / dt;
  }
  const double ip = c.g.ipw[4][i] * s.wind_w[i] * wk0 + c.g.ipw[5][i] * s.wind_w[ib] * wk1;
  c.w.heating_diss[i] = -rho_sum(c, s, i) * ip;
}

// ---------------------------------------------------------------------------------------------
// explicit wind tendencies (mo_vector_tend_expl.F90:128-180) and the forward update of u, v
// (mo_manage_pchevi.F90:86-89)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(BX * BY) k_vec_tend(const Ctx c, const StateDev old, const StateDev nw, int rk) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (jk > nx + 1 || ji > ny + 1 || jl > nl) return;
  const double new_w = rk == 2 ? 0.5 : 1.0, old_w = 1.0 - new_w;
  const double cur_hor = 0.5 + IMPL_THERMO_WEIGHT, old_hor = 1.0 - cur_hor, cur_ver = 1.0 - IMPL_THERMO_WEIGHT;
  const double dt = c.C.dtime;
  const size_t i = LG_IX(c.G, ji, jk, jl);
  if (ji <= ny) {
    const double t = old_w * c.w.tend_u[i] +
                     new_w * (old_hor * c.w.pold_u[i] - cur_hor * (c.w.pnl_u[i] + c.w.pl_u[i]) - 0.5 * c.w.vsg_x[i] + c.w.pvt_x[i]);
    c.w.tend_u[i] = t;
    nw.wind_u[i] = old.wind_u[i] + dt * t;
  }
  if (jk <= nx) {
    const double t = old_w * c.w.tend_v[i] +
                     new_w * (old_hor * c.w.pold_v[i] - cur_hor * (c.w.pnl_v[i] + c.w.pl_v[i]) - 0.5 * c.w.vsg_y[i] + c.w.pvt_y[i]);
    c.w.tend_v[i] = t;
    nw.wind_v[i] = old.wind_v[i] + dt * t;
  }
  if (ji <= ny && jk <= nx && jl >= 2) {
    double x = -cur_ver * (c.w.pnl_w[i] + c.w.pl_w[i]) - 0.5 * c.w.vsg_z[i] + c.w.pvt_z[i];
    if (c.C.lmoist) {  // calc_p_grad_condensates_v, mo_p_grad.F90:106-120 (identically 0 when dry)
      const size_t ia = LG_IX(c.G, ji, jk, jl - 1);
      const double da = old.rho[c.G.ncond][ia] / rho_sum(c, old, ia) - 1.0;
      const double db = old.rho[c.G.ncond][i] / rho_sum(c, old, i) - 1.0;
      x = x + 0.5 * (da + db) * c.g.gravity_m_v[i];
    }
    c.w.tend_w[i] = old_w * c.w.tend_w[i] + new_w * x;
  }
}

// ---------------------------------------------------------------------------------------------
// scalar_tend_expl, mo_scalar_tend_expl.F90:24-181
// ---------------------------------------------------------------------------------------------

// gas mass flux with the new winds and the theta_v flux at u and v points (copies at nx+1, ny+1)
__global__ void __launch_bounds__(BX * BY) k_scalar_flux(const Ctx c, const StateDev ss, const StateDev nw) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (jk > nx + 1 || ji > ny + 1 || jl > nl) return;
  const bool o3 = c.C.order == 3;
  const double *rg = ss.rho[c.G.ncond];
  if (ji <= ny) {
    const int k = (jk == nx + 1) ? 1 : jk, km = wrapm(k, nx);
    const double f = 0.5 * (AT(rg, ji, km, jl) + AT(rg, ji, k, jl)) * AT(nw.wind_u, ji, k, jl);
    double th;
    if (o3)
      th = AT(c.w.tu[0], ji, k, jl) * f;
    else
      th = 0.5 * ((AT(c.g.theta_v_bg, ji, km, jl) + AT(ss.theta_v_pert, ji, km, jl)) +
                  (AT(c.g.theta_v_bg, ji, k, jl) + AT(ss.theta_v_pert, ji, k, jl))) * f;
    AT(c.w.fu, ji, jk, jl) = f;
    AT(c.w.thu, ji, jk, jl) = th;
  }
  if (jk <= nx) {
    const int j = (ji == ny + 1) ? 1 : ji, jm = wrapm(j, ny);
    const double f = 0.5 * (AT(rg, jm, jk, jl) + AT(rg, j, jk, jl)) * AT(nw.wind_v, j, jk, jl);
    double th;
    if (o3)
      th = AT(c.w.tv[0], j, jk, jl) * f;
    else
      th = 0.5 * ((AT(c.g.theta_v_bg, jm, jk, jl) + AT(ss.theta_v_pert, jm, jk, jl)) +
                  (AT(c.g.theta_v_bg, j, jk, jl) + AT(ss.theta_v_pert, j, jk, jl))) * f;
    AT(c.w.fv, ji, jk, jl) = f;
    AT(c.w.thv, ji, jk, jl) = th;
  }
}

// Tracer fluxes of div_h_tracers (mo_divergence_operators.F90:70-147) with upstream face values
// (scalar_times_vector_h_upstream, mo_multiplications.F90:206-287).  The wind-dependent parts
// (face winds, upwind direction, contravariant corrections) are shared by all tracers of a cell.
struct TracerFaces {
  double ue, uw, vn, vs;   // face winds: east (jk+1), west (jk), v at ji, v at ji+1
  size_t re, rw, rn, rs;   // upwind cell of each face
  double cu, cl;           // contravariant corrections at the upper (jl) and lower (jl+1) interface
};

__device__ __forceinline__ TracerFaces tracer_faces(const Ctx &c, const StateDev &nw, int ji, int jk, int jl) {
  const int nx = c.G.nx, ny = c.G.ny, nflat = c.G.nflat, nl = c.G.nl;
  TracerFaces f;
  // face jk+1 (the copy of face 1 when jk = nx)
  {
    const int k = (jk + 1 == nx + 1) ? 1 : jk + 1;
    f.ue = AT(nw.wind_u, ji, k, jl);
    f.re = f.ue >= 0.0 ? LG_IX(c.G, ji, wrapm(k, nx), jl) : LG_IX(c.G, ji, k, jl);
  }
  {
    f.uw = AT(nw.wind_u, ji, jk, jl);
    f.rw = f.uw >= 0.0 ? LG_IX(c.G, ji, wrapm(jk, nx), jl) : LG_IX(c.G, ji, jk, jl);
  }
  {
    f.vn = AT(nw.wind_v, ji, jk, jl);
    f.rn = f.vn >= 0.0 ? LG_IX(c.G, ji, jk, jl) : LG_IX(c.G, wrapm(ji, ny), jk, jl);
  }
  {
    const int j = (ji + 1 == ny + 1) ? 1 : ji + 1;
    f.vs = AT(nw.wind_v, j, jk, jl);
    f.rs = f.vs >= 0.0 ? LG_IX(c.G, j, jk, jl) : LG_IX(c.G, wrapm(j, ny), jk, jl);
  }
  f.cu = 0.0;
  f.cl = 0.0;
  if (jl == nflat) {
    f.cl = vcc(c, nw.wind_u, nw.wind_v, ji, jk, jl + 1);
  } else if (jl == nl) {
    f.cu = vcc(c, nw.wind_u, nw.wind_v, ji, jk, jl);
  } else if (jl > nflat) {
    f.cu = vcc(c, nw.wind_u, nw.wind_v, ji, jk, jl);
    f.cl = vcc(c, nw.wind_u, nw.wind_v, ji, jk, jl + 1);
  }
  return f;
}

// div_h_tracers of one tracer field r, mo_divergence_operators.F90:94-145
__device__ __forceinline__ double div_tracer(const Ctx &c, const TracerFaces &f, const double *r, int ji, int jk, int jl) {
  const double comp_h = (r[f.re] * f.ue) * AT(c.g.area_x, ji, jk + 1, jl) + (r[f.rn] * f.vn) * AT(c.g.area_y, ji, jk, jl) -
                        (r[f.rw] * f.uw) * AT(c.g.area_x, ji, jk, jl) - (r[f.rs] * f.vs) * AT(c.g.area_y, ji + 1, jk, jl);
  double comp_v = 0.0;
  const int nflat = c.G.nflat, nl = c.G.nl;
  if (jl == nflat) {
    const double dl = f.cl <= 0.0 ? AT(r, ji, jk, jl) : AT(r, ji, jk, jl + 1);
    comp_v = -dl * f.cl * AT(c.g.area_z, ji, jk, jl + 1);
  } else if (jl == nl) {
    const double du = f.cu <= 0.0 ? AT(r, ji, jk, jl - 1) : AT(r, ji, jk, jl);
    comp_v = du * f.cu * AT(c.g.area_z, ji, jk, jl);
  } else if (jl > nflat) {
    const double du = f.cu <= 0.0 ? AT(r, ji, jk, jl - 1) : AT(r, ji, jk, jl);
    const double dl = f.cl <= 0.0 ? AT(r, ji, jk, jl) : AT(r, ji, jk, jl + 1);
    comp_v = du * f.cu * AT(c.g.area_z, ji, jk, jl) - dl * f.cl * AT(c.g.area_z, ji, jk, jl + 1);
  }
  return (comp_h + comp_v) / AT(c.g.volume, ji, jk, jl);
}

// divergences -> tend%rho, tend%rhotheta_v; plus w_free_slip (mo_inner_product.F90:52-75) of
// the new state, which only depends on the already final u, v of this substep
__global__ void __launch_bounds__(BX * BY) k_scalar_div(const Ctx c, const StateDev ss, const StateDev nw, int rk) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, ncond = c.G.ncond;
  if (jk > nx || ji > ny || jl > nl) return;
  const size_t i = LG_IX(c.G, ji, jk, jl);
  // gas (new weight 1 at both substeps)
  c.w.tend_rho[ncond][i] = -div_h(c, c.w.fu, c.w.fv, ji, jk, jl);
  const double heat = c.w.heating_diss[i] + c.w.radiation_tendency[i] + c.w.csh[i];
  c.w.tend_rhotheta_v[i] = -div_h(c, c.w.thu, c.w.thv, ji, jk, jl) +
                           C_D_V * ss.rho[ncond][i] / c_v_mass_weighted_air(c, ss, c.w.temperature[i], i) * heat /
                               (C_D_P * (c.g.exner_bg[i] + ss.exner_pert[i]));
  if (c.G.nc > 1) {
    const double new_w = rk == 2 ? 0.5 : 1.0, old_w = 1.0 - new_w;
    const TracerFaces f = tracer_faces(c, nw, ji, jk, jl);
    for (int jc = 0; jc < c.G.nc; ++jc) {
      if (jc == ncond) continue;
      const double d = div_tracer(c, f, ss.rho[jc], ji, jk, jl);
      c.w.tend_rho[jc][i] = old_w * c.w.tend_rho[jc][i] + new_w * (-d);
    }
  }
  if (c.C.lfreeslip && jl == nl) {
    AT(nw.wind_w, ji, jk, nl + 1) = AT(c.g.ipw[0], ji, jk, nl) * AT(nw.wind_u, ji, jk + 1, nl) * AT(c.g.slope_x, ji, jk + 1, nl) +
                                    AT(c.g.ipw[1], ji, jk, nl) * AT(nw.wind_v, ji, jk, nl) * AT(c.g.slope_y, ji, jk, nl) +
                                    AT(c.g.ipw[2], ji, jk, nl) * AT(nw.wind_u, ji, jk, nl) * AT(c.g.slope_x, ji, jk, nl) +
                                    AT(c.g.ipw[3], ji, jk, nl) * AT(nw.wind_v, ji + 1, jk, nl) * AT(c.g.slope_y, ji + 1, jk, nl);
  }
}

}  // namespace lgame
