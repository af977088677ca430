// lgame_step.hip -- HIP stencil kernels of the HEVI predictor-corrector step for gfx950.
//
// One kernel per fused operator group of pchevi (src/time_stepping/mo_manage_pchevi.F90:27-118).
// Horizontal stencil kernels map threadIdx.x to x (64 consecutive x points per wavefront, the
// contiguous direction of the layout), threadIdx.y to y and blockIdx.z to the level.  The two
// implicit column solvers map one column per lane and keep the Thomas recursion in registers
// with per-level values in coalesced column scratch.
#include "lgame_ops.h"
#include "lgame_step.h"
#include "lgame_batch.h"

namespace lgame {

// ---------------------------------------------------------------------------------------------
// temperature_diagnostics, mo_derived.F90:18-38
// ---------------------------------------------------------------------------------------------
// Also the per-cell factors of this state that later kernels of the step read at several cells
// and in both substeps: the condensate deceleration of the pressure gradient C_D_P*rho_gas/rho_sum
// (mo_p_grad.F90:74-104), and with moisture rho_gas/rho_sum - 1 of calc_p_grad_condensates_v
// (mo_p_grad.F90:106-120) and the sink velocity of cloud water, which depends only on this state
// and the temperature (the generalized-density solver of both substeps reads it).
template <class R>
__device__ __forceinline__ void b_temperature(const CtxT<R> &c, const StateDev &s, const Bid bid) {
  KIDX;
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  const ix_t i = LG_IX(c.G, ji, jk, jl);
  const int ncond = c.G.ncond;
  R t = (c.g.theta_v_bg[i] + R(s.theta_v_pert[i])) * (c.g.exner_bg[i] + R(s.exner_pert[i]));
  if (c.C.lmoist) t = t / (R(1.0) + R(s.rho[ncond + 1][i]) / R(s.rho[ncond][i]) * (R(M_D / M_V) - R(1.0)));
  c.w.temperature[i] = t;
  const R rs = rho_sum(c, s, i), rg = R(s.rho[ncond][i]);
  c.w.decel[i] = R(C_D_P) * rg / rs;
  if (c.C.lmom_h || c.C.lmom_v) c.w.rsum[i] = rs;  // read by the momentum diffusion kernels
  if (c.C.lmoist) {
    c.w.pgc[i] = rg / rs - R(1.0);
    c.w.vfall_cw[i] = v_fall_liquid(c, s, ji, jk, jl, (double)t, CLOUD_WATER_RADIUS);
  }
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_temperature(const CtxT<R> c, const StateDev s) {
  b_temperature(c, s, bid_here());
}

// ---------------------------------------------------------------------------------------------
// theta_v_adv_3rd_order, mo_multiplications.F90:116-204, in three passes.
// field f: 0 = full theta_v (theta_v_bg + theta_v_pert), 1 = theta_v_pert.
// ---------------------------------------------------------------------------------------------
// pass 1a: scalar_placeholder = theta_v_bg + theta_v_pert (f=0) or theta_v_pert (f=1)
// pass 1: u,v_placeholder = grad_hor(theta) with its grad_vert inline, theta being the full
// theta_v of state s (field 0) or its theta_v_pert (field 1), formed at every load
template <class R>
__device__ __forceinline__ void b_theta_grad(const CtxT<R> &c, const StateDev &s, int nfield, const Bid bid) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny;
  if (jk > nx + 1 || ji > ny + 1 || jl > c.G.nl) return;
  const ThetaFull<R> full{c.g.theta_v_bg, Fld<const double>(s.theta_v_pert)};
  const Fld<const double> pert = s.theta_v_pert;
  if (ji <= ny) AT(c.w.tgu[0], ji, jk, jl) = grad_hor_x(c, full, ji, jk, jl);
  if (jk <= nx) AT(c.w.tgv[0], ji, jk, jl) = grad_hor_y(c, full, ji, jk, jl);
  if (nfield > 1) {
    if (ji <= ny) AT(c.w.tgu[1], ji, jk, jl) = grad_hor_x(c, pert, ji, jk, jl);
    if (jk <= nx) AT(c.w.tgv[1], ji, jk, jl) = grad_hor_y(c, pert, ji, jk, jl);
  }
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_theta_grad(const CtxT<R> c, const StateDev s, int nfield) {
  b_theta_grad(c, s, nfield, bid_here());
}

// pass 2: flux_density_div = div_h(u_placeholder, v_placeholder)
template <class R>
__device__ __forceinline__ void b_theta_div(const CtxT<R> &c, int nfield, const Bid bid) {
  KIDX;
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  for (int f = 0; f < nfield; ++f) AT(c.w.tfd[f], ji, jk, jl) = div_h(c, c.w.tgu[f], c.w.tgv[f], ji, jk, jl);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_theta_div(const CtxT<R> c, int nfield) {
  b_theta_div(c, nfield, bid_here());
}

// pass 3: upwind-corrected edge values; the upwind direction comes from the given winds
template <class R>
__device__ __forceinline__ void b_theta_edge(const CtxT<R> &c, const StateDev &s, int nfield, const Bid bid) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny;
  if (jk > nx + 1 || ji > ny + 1 || jl > c.G.nl) return;
  const Fld<const double> wind_u = s.wind_u, wind_v = s.wind_v;
  const R c8 = R(1.0) / R(8.0);
  const ThetaFull<R> full{c.g.theta_v_bg, Fld<const double>(s.theta_v_pert)};
  const Fld<const double> pert = s.theta_v_pert;
  for (int f = 0; f < nfield; ++f) {
    // theta at (a, b): the full theta_v (field 0) or theta_v_pert (field 1)
    auto th = [&](int a, int b) { return f == 0 ? ldv(full, c.G, a, b, jl) : R(AT(pert, a, b, jl)); };
    const Fld<const R> fd = c.w.tfd[f];
    if (ji <= ny) {
      const int k = (jk == nx + 1) ? 1 : jk;  // theta_v_u(:,nx+1,:) = theta_v_u(:,1,:)
      const int km = wrapm(k, nx);
      R t = R(0.5) * (th(ji, km) + th(ji, k));
      if (ATR(wind_u, ji, k, jl) >= R(0.0))
        t = t - c8 * R(c.C.dyn) * R(c.C.dxn) * AT(fd, ji, km, jl);
      else
        t = t - c8 * R(c.C.dyn) * R(c.C.dxn) * AT(fd, ji, k, jl);
      AT(c.w.tu[f], ji, jk, jl) = t;
    }
    if (jk <= nx) {
      const int j = (ji == ny + 1) ? 1 : ji;
      const int jm = wrapm(j, ny);
      R t = R(0.5) * (th(j, jk) + th(jm, jk));
      if (ATR(wind_v, j, jk, jl) >= R(0.0))
        t = t - c8 * R(c.C.dyn) * R(c.C.dxn) * AT(fd, j, jk, jl);
      else
        t = t - c8 * R(c.C.dyn) * R(c.C.dxn) * AT(fd, jm, jk, jl);
      AT(c.w.tv[f], ji, jk, jl) = t;
    }
  }
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_theta_edge(const CtxT<R> c, const StateDev s, int nfield) {
  b_theta_edge(c, s, nfield, bid_here());
}

// ---------------------------------------------------------------------------------------------
// manage_p_grad, mo_p_grad.F90:20-104 (luse_bg_state): nonlinear and linear pressure gradient
// accelerations at u, v and w points, scaled by the condensate deceleration factor.
// ---------------------------------------------------------------------------------------------
// The pressure-gradient accelerations at the u, v and w points of one cell (ji <= ny, jk <= nx). Its west, south and upper neighbours and the cell itself are shared by the three
// points, so every state value, metric and deceleration factor is loaded or computed once and
// then combined with the expressions of grad_hor_x/_y, remap_ver2hor_x/_y and the w-point
// gradient, in their order.  Returns the values it stores at the cell's three points.
template <class R>
struct PGrad {
  R nl_u, l_u, old_u;  // u point: nonlinear and linear parts, and the previous step's sum (pold)
  R nl_v, l_v, old_v;  // v point
  R nl_w, l_w;         // w point (levels 2..n_layers)
};
template <class R>
__device__ __forceinline__ PGrad<R> pgrad_interior(const CtxT<R> &c, const StateDev &s, int lfirst, int ji, int jk, int jl) {
  PGrad<R> o;
  o.nl_w = o.l_w = R(0.0);
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev, nflat = c.G.nflat;
  const bool o3 = c.C.order == 3;
  const Geom &G = c.G;
  const ix_t iC = LG_IX(G, ji, jk, jl), iW = LG_IX(G, ji, wrapm(jk, nx), jl), iS = LG_IX(G, wrapm(ji, ny), jk, jl);
  const ix_t dl = (ix_t)G.lvl;  // one level down
  const Fld<const double> ep = s.exner_pert, tp = s.theta_v_pert;
  const bool w_pt = jl >= 2;
  const bool terr = jl > nflat;
  const bool below = jl < nl || c.C.lfreeslip;  // the remap reads the interface below
  // Exner pressure and the vertical gradient of it (grad_vert, zero at levels 1 and n_levels)
  const R eC = R(ep[iC]), eW = R(ep[iW]), eS = R(ep[iS]);
  auto vgrad = [&](ix_t ix, R e_here, int l) {  // level l of the column through ix (at level jl)
    if (l >= nlev || l <= 1) return R(0.0);
    const ix_t il = ix + (l - jl) * dl;
    const R up = (l == jl) ? R(ep[il - dl]) : e_here;
    const R lo = (l == jl) ? e_here : R(ep[il]);
    return (up - lo) / c.g.dz[il];
  };
  R vC = R(0.0), vW = R(0.0), vS = R(0.0), vC1 = R(0.0), vW1 = R(0.0), vS1 = R(0.0);
  if (terr || w_pt) vC = vgrad(iC, eC, jl);
  if (terr) {
    vW = vgrad(iW, eW, jl);
    vS = vgrad(iS, eS, jl);
    if (below) {
      vC1 = vgrad(iC, eC, jl + 1);
      vW1 = vgrad(iW, eW, jl + 1);
      vS1 = vgrad(iS, eS, jl + 1);
    }
  }
  // theta_v, full and perturbation, and the deceleration factors
  // C_D_P*rho_gas/rho_sum of a cell, formed with the temperature at the start of the step
  auto decel_at = [&](ix_t ix) { return c.w.decel[ix]; };
  const R tpC = R(tp[iC]), thC = c.g.theta_v_bg[iC] + tpC, dC = decel_at(iC);
  // u point
  {
    R pold = R(0.0);
    if (!lfirst) {
      pold = -c.w.pnl_u[iC] - R(1.0) * c.w.pl_u[iC];
      c.w.pold_u[iC] = pold;
    }
    R g = (eC - eW) / c.g.dx[iC];
    if (terr) {
      R r = R(0.0);
      r = r + c.g.ipw[4][iW] * vW;
      r = r + c.g.ipw[4][iC] * vC;
      if (below) {
        r = r + c.g.ipw[5][iW] * vW1;
        r = r + c.g.ipw[5][iC] * vC1;
      }
      g = g - c.g.slope_x[iC] * (R(0.5) * r);
    }
    const R tpW = R(tp[iW]);
    R t, lu;
    if (o3) {
      t = c.w.tu[0][iC] * g;
      lu = c.w.tu[1][iC] * c.g.ebg_u[iC];
    } else {
      t = R(0.5) * ((c.g.theta_v_bg[iW] + tpW) + thC) * g;
      lu = R(0.5) * (tpW + tpC) * c.g.ebg_u[iC];
    }
    const R d = R(0.5) * (decel_at(iW) + dC);
    const R nlu = d * t, llu = d * lu;
    c.w.pnl_u[iC] = nlu;
    c.w.pl_u[iC] = llu;
    if (lfirst) {
      pold = -nlu - R(1.0) * llu;
      c.w.pold_u[iC] = pold;
    }
    o.nl_u = nlu;
    o.l_u = llu;
    o.old_u = pold;
    if (jk == 1) {  // column nx+1 is the copy of column 1
      const ix_t iX = LG_IX(G, ji, nx + 1, jl);
      c.w.pnl_u[iX] = nlu;
      c.w.pl_u[iX] = llu;
      c.w.pold_u[iX] = pold;
    }
  }
  // v point
  {
    R pold = R(0.0);
    if (!lfirst) {
      pold = -c.w.pnl_v[iC] - R(1.0) * c.w.pl_v[iC];
      c.w.pold_v[iC] = pold;
    }
    R g = (eS - eC) / c.g.dy[iC];
    if (terr) {
      R r = R(0.0);
      r = r + c.g.ipw[4][iS] * vS;
      r = r + c.g.ipw[4][iC] * vC;
      if (below) {
        r = r + c.g.ipw[5][iS] * vS1;
        r = r + c.g.ipw[5][iC] * vC1;
      }
      g = g - c.g.slope_y[iC] * (R(0.5) * r);
    }
    const R tpS = R(tp[iS]);
    R t, lv;
    if (o3) {
      t = c.w.tv[0][iC] * g;
      lv = c.w.tv[1][iC] * c.g.ebg_v[iC];
    } else {
      t = R(0.5) * ((c.g.theta_v_bg[iS] + tpS) + thC) * g;
      lv = R(0.5) * (tpS + tpC) * c.g.ebg_v[iC];
    }
    const R d = R(0.5) * (decel_at(iS) + dC);
    const R nlv = d * t, llv = d * lv;
    c.w.pnl_v[iC] = nlv;
    c.w.pl_v[iC] = llv;
    if (lfirst) {
      pold = -nlv - R(1.0) * llv;
      c.w.pold_v[iC] = pold;
    }
    o.nl_v = nlv;
    o.l_v = llv;
    o.old_v = pold;
    if (ji == 1) {  // row ny+1 is the copy of row 1
      const ix_t iY = LG_IX(G, ny + 1, jk, jl);
      c.w.pnl_v[iY] = nlv;
      c.w.pl_v[iY] = llv;
      c.w.pold_v[iY] = pold;
    }
  }
  // w point, levels 2..n_layers
  if (w_pt) {
    const ix_t iA = iC - dl;
    const R tpA = R(tp[iA]);
    const R t = R(0.5) * ((c.g.theta_v_bg[iA] + tpA) + thC) * vC;
    const R lw = R(0.5) * (tpA + tpC) * c.g.ebg_w[iC];
    const R d = R(0.5) * (decel_at(iA) + dC);
    o.nl_w = d * t;
    o.l_w = d * lw;
    c.w.pnl_w[iC] = o.nl_w;
    c.w.pl_w[iC] = o.l_w;
  }
  return o;
}


// ---------------------------------------------------------------------------------------------
// Momentum advection block of vector_tend_expl (mo_vector_tend_expl.F90:58-99):
// potential vorticity, mass fluxes, vorticity flux term, kinetic-energy gradient.
// ---------------------------------------------------------------------------------------------

// circulation of zeta_x around its dual face, mo_vorticities.F90:33-93 (jl in 2..n_levels)
template <class R, class X>
__device__ VT<R, X> zeta_x_circ(const CtxT<R> &c, const StateDev &s, int ji, X jk, int jl) {
  const int ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev;
  const int jm = wrapm(ji, ny);
  if (jl <= nl) {
    return LDG(c.g.dz, jm, jk, jl) * LDGR(s.wind_w, jm, jk, jl) - LDG(c.g.dy, ji, jk, jl - 1) * hcov_y(c, s.wind_v, s.wind_w, ji, jk, jl - 1) -
           LDG(c.g.dz, ji, jk, jl) * LDGR(s.wind_w, ji, jk, jl) + LDG(c.g.dy, ji, jk, jl) * hcov_y(c, s.wind_v, s.wind_w, ji, jk, jl);
  }
  VT<R, X> r = -LDG(c.g.dy, ji, jk, nl) * hcov_y(c, s.wind_v, s.wind_w, ji, jk, nl);
  if (c.C.lfreeslip) {
    r = r + R(0.5) * LDG(c.g.dz, jm, jk, nlev) * LDGR(s.wind_w, jm, jk, nlev) - R(0.5) * LDG(c.g.dz, ji, jk, nlev) * LDGR(s.wind_w, ji, jk, nlev) +
        LDG(c.g.dy, ji, jk, nl) *
            (LDGR(s.wind_v, ji, jk, nl) +
             (LDGR(s.wind_v, ji, jk, nl - 1) - LDGR(s.wind_v, ji, jk, nl)) / (LDG(c.g.z_v, ji, jk, nl - 1) - LDG(c.g.z_v, ji, jk, nl)) *
                 (R(0.5) * (LDG(c.g.z_w, jm, jk, nlev) + LDG(c.g.z_w, ji, jk, nlev)) - LDG(c.g.z_v, ji, jk, nl)) +
             R(0.5) * (LDGR(s.wind_w, jm, jk, nlev) + LDGR(s.wind_w, ji, jk, nlev)) * (LDG(c.g.z_w, jm, jk, nlev) - LDG(c.g.z_w, ji, jk, nlev)) /
                 LDG(c.g.dy, ji, jk, nl));
  }
  return r;
}

// mo_vorticities.F90:105-165 (jl in 2..n_levels)
template <class R, class X>
__device__ VT<R, X> zeta_y_circ(const CtxT<R> &c, const StateDev &s, int ji, X jk, int jl) {
  const int nx = c.G.nx, nl = c.G.nl, nlev = c.G.nlev;
  const auto km = wrapm(jk, nx);
  if (jl <= nl) {
    return -LDG(c.g.dz, ji, jk, jl) * LDGR(s.wind_w, ji, jk, jl) - LDG(c.g.dx, ji, jk, jl) * hcov_x(c, s.wind_u, s.wind_w, ji, jk, jl) +
           LDG(c.g.dz, ji, km, jl) * LDGR(s.wind_w, ji, km, jl) + LDG(c.g.dx, ji, jk, jl - 1) * hcov_x(c, s.wind_u, s.wind_w, ji, jk, jl - 1);
  }
  VT<R, X> r = LDG(c.g.dx, ji, jk, nl) * hcov_x(c, s.wind_u, s.wind_w, ji, jk, nl);
  if (c.C.lfreeslip) {
    r = r + R(0.5) * LDG(c.g.dz, ji, km, nlev) * LDGR(s.wind_w, ji, km, nlev) - R(0.5) * LDG(c.g.dz, ji, jk, nlev) * LDGR(s.wind_w, ji, jk, nlev) -
        LDG(c.g.dx, ji, jk, nl) *
            (LDGR(s.wind_u, ji, jk, nl) +
             (LDGR(s.wind_u, ji, jk, nl - 1) - LDGR(s.wind_u, ji, jk, nl)) / (LDG(c.g.z_u, ji, jk, nl - 1) - LDG(c.g.z_u, ji, jk, nl)) *
                 (R(0.5) * (LDG(c.g.z_w, ji, km, nlev) + LDG(c.g.z_w, ji, jk, nlev)) - LDG(c.g.z_u, ji, jk, nl)) +
             R(0.5) * (LDGR(s.wind_w, ji, km, nlev) + LDGR(s.wind_w, ji, jk, nlev)) * (LDG(c.g.z_w, ji, jk, nlev) - LDG(c.g.z_w, ji, km, nlev)) /
                 LDG(c.g.dx, ji, jk, nl));
  }
  return r;
}

// density average of calc_pot_vort for horizontal vorticities, mo_vorticities.F90:240-266
template <class R, class XA, class XB>
__device__ VT<R, XA> pv_den_h(const CtxT<R> &c, const StateDev &s, int ia, XA ka, int ib, XB kb, int jl) {
  if (jl == 1) {
    return R(0.5) * (rhog(c, s, ia, ka, jl) +
                  (R(c.C.toa) - LDG(c.g.z_scalar, ia, ka, jl)) * (rhog(c, s, ia, ka, jl) - rhog(c, s, ia, ka, jl + 1)) /
                      (LDG(c.g.z_scalar, ia, ka, jl) - LDG(c.g.z_scalar, ia, ka, jl + 1)) +
                  rhog(c, s, ib, kb, jl) +
                  (R(c.C.toa) - LDG(c.g.z_scalar, ib, kb, jl)) * (rhog(c, s, ib, kb, jl) - rhog(c, s, ib, kb, jl + 1)) /
                      (LDG(c.g.z_scalar, ib, kb, jl) - LDG(c.g.z_scalar, ib, kb, jl + 1)));
  } else if (jl == c.G.nlev) {
    return R(0.5) * (rhog(c, s, ia, ka, jl - 1) +
                  (LDG(c.g.z_w, ia, ka, jl) - LDG(c.g.z_scalar, ia, ka, jl - 1)) * (rhog(c, s, ia, ka, jl - 2) - rhog(c, s, ia, ka, jl - 1)) /
                      (LDG(c.g.z_scalar, ia, ka, jl - 2) - LDG(c.g.z_scalar, ia, ka, jl - 1)) +
                  rhog(c, s, ib, kb, jl - 1) +
                  (LDG(c.g.z_w, ib, kb, jl) - LDG(c.g.z_scalar, ib, kb, jl - 1)) * (rhog(c, s, ib, kb, jl - 2) - rhog(c, s, ib, kb, jl - 1)) /
                      (LDG(c.g.z_scalar, ib, kb, jl - 2) - LDG(c.g.z_scalar, ib, kb, jl - 1)));
  }
  return R(0.25) * (rhog(c, s, ia, ka, jl - 1) + rhog(c, s, ib, kb, jl - 1) + rhog(c, s, ia, ka, jl) + rhog(c, s, ib, kb, jl));
}

// calc_pot_vort (mo_vorticities.F90:189-421) with rel_vort (:19-187) inline: the three potential
// vorticities of cell (ji, jk, jl) (of two cells for X = XP), formed before any store so that the
// loads they share are made once.  eta_z is zero above the layers.
template <class R, class X>
struct PotVort {
  VT<R, X> ex, ey, ez;
};
template <class R, class X>
__device__ __forceinline__ PotVort<R, X> pot_vort_at(const CtxT<R> &c, const StateDev &s, int ji, X jk, int jl) {
  using V = VT<R, X>;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  const int jm = wrapm(ji, ny);
  const auto km = wrapm(jk, nx);
  const int le = jl == 1 ? 2 : jl;  // zeta_x,y(:,:,1) = zeta_x,y(:,:,2)
  PotVort<R, X> o;
  // eta_x
  V zx = R(0.0);
  if (!c.C.llinear) zx = zeta_x_circ(c, s, ji, jk, le) / LDG(c.g.area_dual_x, ji, jk, le);
  o.ex = (c.C.lcorio ? zx + LDG(c.g.fvec_x, ji, jk, 1) : zx) / pv_den_h(c, s, jm, jk, ji, jk, jl);
  // eta_y
  V zy = R(0.0);
  if (!c.C.llinear) zy = zeta_y_circ(c, s, ji, jk, le) / LDG(c.g.area_dual_y, ji, jk, le);
  o.ey = (c.C.lcorio ? zy + LDG(c.g.fvec_y, ji, jk, 1) : zy) / pv_den_h(c, s, ji, km, ji, jk, jl);
  // eta_z
  o.ez = R(0.0);
  if (jl <= nl) {
    V z = R(0.0);
    if (!c.C.llinear) z = rel_vort_z(c, s, ji, jk, jl) / LDG(c.g.area_dual_z, ji, jk, jl);
    const V e = c.C.lcorio ? z + LDG(c.g.fvec_z, ji, jk, 1) : z;
    o.ez = e / (R(0.25) * (rhog(c, s, ji, km, jl) + rhog(c, s, ji, jk, jl) + rhog(c, s, jm, jk, jl) + rhog(c, s, jm, km, jl)));
  }
  return o;
}
// mass flux densities of the gas, scalar_times_vector_h/v (mo_vector_tend_expl.F90:64-80), and the
// kinetic energy (inner_product, mo_inner_product.F90:16-50) of cell (ji, jk, jl), or of the two
// cells of a column pair, formed from the same density and wind values as the potential
// vorticity
template <class R, class X>
struct MassFlux {
  VT<R, X> fu, fv, fw, ke;
};
template <class R, class X>
__device__ __forceinline__ MassFlux<R, X> mass_flux_at(const CtxT<R> &c, const StateDev &s, int ji, X jk, int jl, int with_ke) {
  using V = VT<R, X>;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev;
  MassFlux<R, X> o;
  o.ke = R(0.0);
  if (with_ke && jl <= nl) o.ke = ke_point(c, s, ji, jk, jl);
  const auto km = wrapm(jk, nx);
  const int jm = wrapm(ji, ny);
  o.fu = o.fv = o.fw = R(0.0);
  if (jl <= nl) {
    o.fu = R(0.5) * (rhog(c, s, ji, km, jl) + rhog(c, s, ji, jk, jl)) * LDGR(s.wind_u, ji, jk, jl);
    o.fv = R(0.5) * (rhog(c, s, jm, jk, jl) + rhog(c, s, ji, jk, jl)) * LDGR(s.wind_v, ji, jk, jl);
  }
  if (jl >= 2 && jl <= nl) {
    o.fw = R(0.5) * (rhog(c, s, ji, jk, jl - 1) + rhog(c, s, ji, jk, jl)) * LDGR(s.wind_w, ji, jk, jl);
  } else if (jl == nlev && c.C.lfreeslip) {
    const V dv = rhog(c, s, ji, jk, nl) + (rhog(c, s, ji, jk, nl - 1) - rhog(c, s, ji, jk, nl)) /
                                              (LDG(c.g.z_scalar, ji, jk, nl - 1) - LDG(c.g.z_scalar, ji, jk, nl)) *
                                              (LDG(c.g.z_w, ji, jk, nlev) - LDG(c.g.z_scalar, ji, jk, nl));
    o.fw = dv * LDGR(s.wind_w, ji, jk, nlev);
  }
  return o;
}
// The u points of column nx+1 and the v points of row ny+1 are copies of column 1 and row 1.
template <class R, class X>
__device__ __forceinline__ void store_mass_flux(const CtxT<R> &c, int ji, X jk, int jl, int with_ke, const MassFlux<R, X> &m) {
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (jl <= nl) {
    if (with_ke) stv(c.w.v_squared, c.G, ji, jk, jl, m.ke);
    stv(c.w.u_ph, c.G, ji, jk, jl, m.fu);
    stv(c.w.v_ph, c.G, ji, jk, jl, m.fv);
    if (col1(jk)) AT(c.w.u_ph, ji, nx + 1, jl) = first(m.fu);
    if (ji == 1) stv(c.w.v_ph, c.G, ny + 1, jk, jl, m.fv);
  }
  stv(c.w.w_ph, c.G, ji, jk, jl, m.fw);
}

// The values of row ny+1 and column nx+1 are copies of row 1 and column 1 (the reference
// evaluates them with the indices of row 1 and column 1): the threads of row 1 and column 1
// store them, and the launch covers the nx by ny cells.
// The same kernels form the mass fluxes and (with_ke) the kinetic energy of the cells.
template <class R>
__device__ __forceinline__ void b_pot_vort(const CtxT<R> &c, const StateDev &s, int with_ke, const Bid bid) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nlev);
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev;
  if (jk > nx || ji > ny || jl > nlev) return;
  const PotVort<R, int> o = pot_vort_at(c, s, ji, jk, jl);
  const MassFlux<R, int> m = mass_flux_at(c, s, ji, jk, jl, with_ke);
  store_mass_flux(c, ji, jk, jl, with_ke, m);
  AT(c.w.eta_x, ji, jk, jl) = o.ex;
  AT(c.w.eta_y, ji, jk, jl) = o.ey;
  if (jl <= nl) AT(c.w.eta_z, ji, jk, jl) = o.ez;
  if (ji == 1) {
    AT(c.w.eta_x, ny + 1, jk, jl) = o.ex;
    if (jl <= nl) AT(c.w.eta_z, ny + 1, jk, jl) = o.ez;
  }
  if (jk == 1) {
    AT(c.w.eta_y, ji, nx + 1, jl) = o.ey;
    if (jl <= nl) AT(c.w.eta_z, ji, nx + 1, jl) = o.ez;
  }
  if (ji == 1 && jk == 1 && jl <= nl) AT(c.w.eta_z, ny + 1, nx + 1, jl) = o.ez;
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_pot_vort(const CtxT<R> c, const StateDev s, int with_ke) {
  b_pot_vort(c, s, with_ke, bid_here());
}
// The same with two x-adjacent cells per thread (nx even): the launch covers nx/2 by ny pairs.
template <class R>
__global__ void __launch_bounds__(BX * BY) k_pot_vort2(const CtxT<R> c, const StateDev s, int with_ke) {
  const Bid bid = bid_here();
  KIDX_ZFAST(c.G.nx / 2, c.G.ny, c.G.nlev);
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl, nlev = c.G.nlev;
  const int k = 2 * jk - 1;  // the pair's first column
  if (k > nx || ji > ny || jl > nlev) return;
  __builtin_assume(k >= 1 && k < nx);
  const XP x{k};
  const PotVort<R, XP> o = pot_vort_at(c, s, ji, x, jl);
  const MassFlux<R, XP> m = mass_flux_at(c, s, ji, x, jl, with_ke);
  store_mass_flux(c, ji, x, jl, with_ke, m);
  stv(c.w.eta_x, c.G, ji, x, jl, o.ex);
  stv(c.w.eta_y, c.G, ji, x, jl, o.ey);
  if (jl <= nl) stv(c.w.eta_z, c.G, ji, x, jl, o.ez);
  if (ji == 1) {
    stv(c.w.eta_x, c.G, ny + 1, x, jl, o.ex);
    if (jl <= nl) stv(c.w.eta_z, c.G, ny + 1, x, jl, o.ez);
  }
  if (k == 1) {
    AT(c.w.eta_y, ji, nx + 1, jl) = o.ey.a;
    if (jl <= nl) AT(c.w.eta_z, ji, nx + 1, jl) = o.ez.a;
  }
  if (ji == 1 && k == 1 && jl <= nl) AT(c.w.eta_z, ny + 1, nx + 1, jl) = o.ez.a;
}

// The vorticity flux term at the u, v and w points of one cell (ji <= ny, jk <= nx): every eta,
// mass-flux and weight value that two of the points share is loaded once, and the three results
// are stored after all loads.  The expressions and their order are those of the reference's
// per-point loops.  Returns the three values.
template <class R>
struct VortFlux {
  R x, y, z;  // at the u, v and w points (z is zero at level 1)
};
template <class R>
__device__ __forceinline__ VortFlux<R> vort_flux_interior(const CtxT<R> &c, int ji, int jk, int jl) {
  const Geom &G = c.G;
  const int nx = G.nx, ny = G.ny;
  const Fld<const R> ez = c.w.eta_z, up = c.w.u_ph, vp = c.w.v_ph, wp = c.w.w_ph;
  const Fld<const R> ex = c.w.eta_x, ey = c.w.eta_y;
  const int km = wrapm(jk, nx), kp = jk + 1, jm = wrapm(ji, ny);
  const ix_t dl = (ix_t)G.lvl;
  const ix_t iC = LG_IX(G, ji, jk, jl), iW = LG_IX(G, ji, km, jl), iE = LG_IX(G, ji, kp, jl);
  const ix_t iN = LG_IX(G, ji + 1, jk, jl), iNW = LG_IX(G, ji + 1, km, jl), iNE = LG_IX(G, ji + 1, kp, jl);
  const ix_t iS = LG_IX(G, jm, jk, jl), iSE = LG_IX(G, jm, kp, jl);
  // eta_z at the eight dual points around the cell
  const R zC = ez[iC], zW = ez[iW], zE = ez[iE], zN = ez[iN], zNW = ez[iNW], zNE = ez[iNE], zS = ez[iS], zSE = ez[iSE];
  const R uW = up[iW], uC = up[iC], uE = up[iE], uS = up[iS], uSE = up[iSE];
  const R vW = vp[iW], vNW = vp[iNW], vN = vp[iN], vC = vp[iC];
  const R wW0 = wp[iW], wW1 = wp[iW + dl], wC0 = wp[iC], wC1 = wp[iC + dl], wS0 = wp[iS], wS1 = wp[iS + dl];
  const R ey0 = ey[iC], ey1 = ey[iC + dl], ex0 = ex[iC], ex1 = ex[iC + dl];
  const R p4W = c.g.ipw[4][iW], p5W = c.g.ipw[5][iW], p4C = c.g.ipw[4][iC], p5C = c.g.ipw[5][iC];
  const R p4S = c.g.ipw[4][iS], p5S = c.g.ipw[5][iS];
  // u point (ji, jk)
  const R *tu = c.g.trsk_u + 3 * (ji - 1);
  R ru = tu[0] * vW * R(0.25) * (zW + zC + zC + zN) + tu[1] * uW * R(0.25) * (zW + zNW + zC + zN) +
         tu[2] * vNW * R(0.25) * (zNW + zN + zC + zN) + tu[2] * vN * R(0.25) * (zN + zNE + zC + zN) +
         tu[1] * uE * R(0.25) * (zNE + zE + zC + zN) + tu[0] * vC * R(0.25) * (zC + zE + zC + zN);
  ru = ru - R(0.5) * p4W * wW0 * ey0 - R(0.5) * p5W * wW1 * ey1 - R(0.5) * p4C * wC0 * ey0 - R(0.5) * p5C * wC1 * ey1;
  // v point (ji, jk)
  const R *tv = c.g.trsk_v + 2 * (ji - 1);
  R rv = tv[0] * uC * R(0.25) * (zC + zN + zC + zE) + tv[0] * uE * R(0.25) * (zE + zNE + zC + zE) +
         tv[1] * uSE * R(0.25) * (zSE + zE + zC + zE) + tv[1] * uS * R(0.25) * (zS + zC + zC + zE);
  rv = rv + R(0.5) * p4S * wS0 * ex0 + R(0.5) * p5S * wS1 * ex1 + R(0.5) * p4C * wC0 * ex0 + R(0.5) * p5C * wC1 * ex1;
  // w point, levels 2..n_layers
  R rw = R(0.0);
  if (jl >= 2) {
    const ix_t a = iC - dl;
    rw = R(0.5) * (c.g.ipw[0][a] * up[iE - dl] * ey[iE] - c.g.ipw[1][a] * vp[a] * ex0 + c.g.ipw[2][a] * up[a] * ey0 -
                   c.g.ipw[3][a] * vp[iN - dl] * ex[iN] + c.g.ipw[0][iC] * uE * ey[iE] - c.g.ipw[1][iC] * vC * ex0 +
                   c.g.ipw[2][iC] * uC * ey0 - c.g.ipw[3][iC] * vN * ex[iN]);
  }
  c.w.pvt_x[iC] = ru;
  c.w.pvt_y[iC] = rv;
  if (jl >= 2) c.w.pvt_z[iC] = rw;
  if (jk == 1) c.w.pvt_x[LG_IX(G, ji, nx + 1, jl)] = ru;
  if (ji == 1) c.w.pvt_y[LG_IX(G, ny + 1, jk, jl)] = rv;
  return VortFlux<R>{ru, rv, rw};
}

// calc_vorticity_flux_term, mo_vorticity_flux.F90:16-171.  The u points of column nx+1 and the
// v points of row ny+1 are copies of column 1 and row 1 (the reference evaluates them with the
// indices of column 1 and row 1), so the threads of column 1 and row 1 store them too and the
// launch covers the nx by ny cells only.
template <class R>
__device__ __forceinline__ void b_vort_flux(const CtxT<R> &c, int with_ke_grad, const Bid bid) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  if (jk > c.G.nx || ji > c.G.ny || jl > c.G.nl) return;
  vort_flux_interior(c, ji, jk, jl);
  // grad_vert + grad_hor of the kinetic energy (mo_vector_tend_expl.F90:96-97), the same cell
  if (with_ke_grad) ke_grad_interior(c, ji, jk, jl);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_vort_flux(const CtxT<R> c, int with_ke_grad) {
  b_vort_flux(c, with_ke_grad, bid_here());
}


// grad_hor and grad_vert of the kinetic energy at the u, v and w points of one cell (ji <= ny,
// jk <= nx): the values and vertical gradients that the three share are loaded or formed once and
// combined with the expressions of grad_hor_x/_y and remap_ver2hor_x/_y, in their order.
template <class R>
struct KeGrad {
  R x, y, z;  // at the u, v and w points (z is zero at level 1)
  R x_edge;   // at the u point of column nx+1 (thread of column 1)
  R y_edge;   // at the v point of row ny+1 (thread of row 1)
};
template <class R>
__device__ __forceinline__ KeGrad<R> ke_grad_interior(const CtxT<R> &c, int ji, int jk, int jl) {
  const Geom &G = c.G;
  const int nx = G.nx, ny = G.ny, nl = G.nl, nlev = G.nlev, nflat = G.nflat;
  const Fld<const R> k2 = c.w.v_squared;
  const ix_t iC = LG_IX(G, ji, jk, jl), iW = LG_IX(G, ji, wrapm(jk, nx), jl), iS = LG_IX(G, wrapm(ji, ny), jk, jl);
  const ix_t dl = (ix_t)G.lvl;
  const bool terr = jl > nflat, below = jl < nl || c.C.lfreeslip;
  const R kC = k2[iC], kW = k2[iW], kS = k2[iS];
  auto vgrad = [&](ix_t ix, R k_here, int l) {  // grad_vert at level l of the column through ix
    if (l >= nlev || l <= 1) return R(0.0);
    const ix_t il = ix + (l - jl) * dl;
    const R up = (l == jl) ? k2[il - dl] : k_here;
    const R lo = (l == jl) ? k_here : k2[il];
    return (up - lo) / c.g.dz[il];
  };
  R vC = R(0.0), vW = R(0.0), vS = R(0.0), vC1 = R(0.0), vW1 = R(0.0), vS1 = R(0.0);
  if (terr || jl >= 2) vC = vgrad(iC, kC, jl);
  if (terr) {
    vW = vgrad(iW, kW, jl);
    vS = vgrad(iS, kS, jl);
    if (below) {
      vC1 = vgrad(iC, kC, jl + 1);
      vW1 = vgrad(iW, kW, jl + 1);
      vS1 = vgrad(iS, kS, jl + 1);
    }
  }
  const R gx_cov = (kC - kW) / c.g.dx[iC];
  const R gy_cov = (kS - kC) / c.g.dy[iC];
  R gx = gx_cov, gy = gy_cov, rx = R(0.0), ry = R(0.0);
  if (terr) {
    rx = rx + c.g.ipw[4][iW] * vW;
    rx = rx + c.g.ipw[4][iC] * vC;
    if (below) {
      rx = rx + c.g.ipw[5][iW] * vW1;
      rx = rx + c.g.ipw[5][iC] * vC1;
    }
    gx = gx - c.g.slope_x[iC] * (R(0.5) * rx);
    ry = ry + c.g.ipw[4][iS] * vS;
    ry = ry + c.g.ipw[4][iC] * vC;
    if (below) {
      ry = ry + c.g.ipw[5][iS] * vS1;
      ry = ry + c.g.ipw[5][iC] * vC1;
    }
    gy = gy - c.g.slope_y[iC] * (R(0.5) * ry);
  }
  KeGrad<R> o{gx, gy, jl >= 2 ? vC : R(0.0), R(0.0), R(0.0)};
  c.w.vsg_x[iC] = gx;
  c.w.vsg_y[iC] = gy;
  if (jl >= 2) c.w.vsg_z[iC] = vC;
  // The u points of column nx+1 and the v points of row ny+1: grad_hor_x/_y there take the
  // covariant part and the remap of column 1 / row 1 (their neighbours wrap to the same cells)
  // and their own slope.
  if (jk == 1) {
    const ix_t iX = LG_IX(G, ji, nx + 1, jl);
    o.x_edge = terr ? gx_cov - c.g.slope_x[iX] * (R(0.5) * rx) : gx_cov;
    c.w.vsg_x[iX] = o.x_edge;
  }
  if (ji == 1) {
    const ix_t iY = LG_IX(G, ny + 1, jk, jl);
    o.y_edge = terr ? gy_cov - c.g.slope_y[iY] * (R(0.5) * ry) : gy_cov;
    c.w.vsg_y[iY] = o.y_edge;
  }
  return o;
}


// simple_dissipation_rate, mo_momentum_diff_diss.F90:333-377, and the vertical diffusion of w
// of mom_diff_v (:277-289,320-328).  mom_diff_tend_z at the two levels of the cell is formed here
// from the coefficient kv of k_diff_coef (vert_vert_mom_viscosity, mo_eff_diff_coeffs.F90:491-524)
// and stored at the upper level; the Klemp layer acceleration is added to it and the inner
// product with the wind gives the heating rate.  mom_diff_v's vertical diffusion of u, v and
// horizontal diffusion of w multiply vert_hor_viscosity_u/_v, which the reference never sets
// (vert_hor_mom_viscosity, :312-489, has no caller; control.F90:427-428 zeroes them): those terms
// are exact zeros and are not formed.
template <class R>
__device__ __forceinline__ void b_dissipation(const CtxT<R> &c, const StateDev &s, const Bid bid) {
  KIDX;
  const Geom &G = c.G;
  const int nl = G.nl;
  if (jk > G.nx || ji > G.ny || jl > nl) return;
  const ix_t i = LG_IX(G, ji, jk, jl), ib = LG_IX(G, ji, jk, jl + 1);
  R mz0 = R(0.0), mz1 = R(0.0);
  if (c.C.lmom_v) {
    // rho*kv*(dw/dz) of layer l: add_vertical_div into a zeroed field, then vert_vert_mom_viscosity
    auto sp = [&](int l) {
      const ix_t q = LG_IX(G, ji, jk, l);
      const R cu = l == 1 ? R(0.0) : R(s.wind_w[q]);
      const R cl = l == nl ? R(0.0) : ATR(s.wind_w, ji, jk, l + 1);
      const R comp_v = cu * c.g.area_z[q] - cl * AT(c.g.area_z, ji, jk, l + 1);
      const R x = R(0.0) + comp_v / c.g.volume[q];
      return R(s.rho[G.ncond][q]) * c.w.kv[q] * x;
    };
    const R sp_c = sp(jl);
    if (jl >= 2) {
      const ix_t ia = LG_IX(G, ji, jk, jl - 1);
      mz0 = (sp(jl - 1) - sp_c) / c.g.dz[i] / (R(0.5) * (c.w.rsum[ia] + c.w.rsum[i]));
      c.w.mom_z[i] = mz0;
    }
    if (jl < nl) mz1 = (sp_c - sp(jl + 1)) / c.g.dz[ib] / (R(0.5) * (c.w.rsum[i] + c.w.rsum[ib]));
  }
  if (!c.C.lklemp && !c.C.lmom_h && !c.C.lpbl) return;  // mo_vector_tend_expl.F90:123-125
  R wk0 = mz0, wk1 = mz1;
  const R dt = R(c.C.dtime);
  if (c.C.lklemp && jl >= 2 && jl <= c.C.n_damping_levels) {
    const R w = R(s.wind_w[i]);
    const R k = (w / c.g.klemp_pref[i] - w) / dt;
    wk0 = c.C.lmom_v ? wk0 + k : k;
  }
  if (c.C.lklemp && jl + 1 <= c.C.n_damping_levels) {
    const R w = R(s.wind_w[ib]);
    const R k = (w / c.g.klemp_pref[ib] - w) / dt;
    wk1 = c.C.lmom_v ? wk1 + k : k;
  }
  R ip;
  if (c.C.lmom_h || c.C.lpbl) {
    ip = AT(c.g.ipw[0], ji, jk, jl) * ATR(s.wind_u, ji, jk + 1, jl) * AT(c.w.mom_x, ji, jk + 1, jl) +
         c.g.ipw[1][i] * R(s.wind_v[i]) * c.w.mom_y[i] + c.g.ipw[2][i] * R(s.wind_u[i]) * c.w.mom_x[i] +
         AT(c.g.ipw[3], ji, jk, jl) * ATR(s.wind_v, ji + 1, jk, jl) * AT(c.w.mom_y, ji + 1, jk, jl) +
         c.g.ipw[4][i] * R(s.wind_w[i]) * wk0 + c.g.ipw[5][i] * R(s.wind_w[ib]) * wk1;
  } else {  // the horizontal terms are products with zero tendencies
    ip = c.g.ipw[4][i] * R(s.wind_w[i]) * wk0 + c.g.ipw[5][i] * R(s.wind_w[ib]) * wk1;
  }
  const R rs = (c.C.lmom_h || c.C.lmom_v) ? c.w.rsum[i] : rho_sum(c, s, i);
  c.w.heating_diss[i] = -rs * ip;
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_dissipation(const CtxT<R> c, const StateDev s) {
  b_dissipation(c, s, bid_here());
}

// ---------------------------------------------------------------------------------------------
// explicit wind tendencies (mo_vector_tend_expl.F90:128-180) and the forward update of u, v
// (mo_manage_pchevi.F90:86-89)
// ---------------------------------------------------------------------------------------------
// the weights of a substep's tendency update
template <class R>
struct VtWeights {
  R new_w, old_w, cur_hor, old_hor, cur_ver;
  bool mom;  // diag%mom_diff_tend_* are added (zero arrays otherwise)
};
template <class R>
__device__ __forceinline__ VtWeights<R> vt_weights(const CtxT<R> &c, int rk) {
  VtWeights<R> w;
  w.new_w = rk == 2 ? R(0.5) : R(1.0);
  w.old_w = R(1.0) - w.new_w;
  w.cur_hor = R(0.5) + R(IMPL_THERMO_WEIGHT);
  w.old_hor = R(1.0) - w.cur_hor;
  w.cur_ver = R(1.0) - R(IMPL_THERMO_WEIGHT);
  w.mom = c.C.lmom_h || c.C.lmom_v || c.C.lpbl;
  return w;
}
// the u point at index i, from its pressure-gradient terms (pold, pnl, pl)
// and the kinetic-energy gradient and vorticity flux term there (vsg, pvt)
template <class R>
__device__ __forceinline__ void vt_u(const CtxT<R> &c, const StateDev &old, const StateDev &nw, const VtWeights<R> &w, ix_t i,
                                     R pold, R pnl, R pl, R vsg, R pvt) {
  R x = w.old_hor * pold - w.cur_hor * (pnl + pl) - R(0.5) * vsg + pvt;
  if (w.mom) x = x + c.w.mom_x[i];
  const R t = w.old_w * c.w.tend_u[i] + w.new_w * x;
  c.w.tend_u[i] = t;
  nw.wind_u[i] = old.wind_u[i] + c.C.dtime * (double)t;  // FP64 state update
}
template <class R>
__device__ __forceinline__ void vt_v(const CtxT<R> &c, const StateDev &old, const StateDev &nw, const VtWeights<R> &w, ix_t i,
                                     R pold, R pnl, R pl, R vsg, R pvt) {
  R x = w.old_hor * pold - w.cur_hor * (pnl + pl) - R(0.5) * vsg + pvt;
  if (w.mom) x = x + c.w.mom_y[i];
  const R t = w.old_w * c.w.tend_v[i] + w.new_w * x;
  c.w.tend_v[i] = t;
  nw.wind_v[i] = old.wind_v[i] + c.C.dtime * (double)t;
}
// the w point of cell (ji, jk, jl), jl >= 2
template <class R>
__device__ __forceinline__ void vt_w(const CtxT<R> &c, const VtWeights<R> &w, int ji, int jk, int jl, ix_t i, R pnl, R pl, R vsg,
                                     R pvt) {
  R x = -w.cur_ver * (pnl + pl) - R(0.5) * vsg + pvt;
  if (c.C.lmoist) {  // calc_p_grad_condensates_v, mo_p_grad.F90:106-120 (identically 0 when dry)
    const ix_t ia = LG_IX(c.G, ji, jk, jl - 1);
    const R da = c.w.pgc[ia];  // rho_gas/rho_sum - 1 of the old state, from k_temperature
    const R db = c.w.pgc[i];
    x = x + R(0.5) * (da + db) * c.g.gravity_m_v[i];
  }
  if (w.mom) x = x + c.w.mom_z[i];
  c.w.tend_w[i] = w.old_w * c.w.tend_w[i] + w.new_w * x;
}
template <class R>
__device__ __forceinline__ void b_vec_tend(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int rk, const Bid bid) {
  KIDX;
  const int nx = c.G.nx, ny = c.G.ny, nl = c.G.nl;
  if (jk > nx + 1 || ji > ny + 1 || jl > nl) return;
  const VtWeights<R> w = vt_weights(c, rk);
  const ix_t i = LG_IX(c.G, ji, jk, jl);
  if (ji <= ny) vt_u(c, old, nw, w, i, c.w.pold_u[i], c.w.pnl_u[i], c.w.pl_u[i], c.w.vsg_x[i], c.w.pvt_x[i]);
  if (jk <= nx) vt_v(c, old, nw, w, i, c.w.pold_v[i], c.w.pnl_v[i], c.w.pl_v[i], c.w.vsg_y[i], c.w.pvt_y[i]);
  if (ji <= ny && jk <= nx && jl >= 2) vt_w(c, w, ji, jk, jl, i, c.w.pnl_w[i], c.w.pl_w[i], c.w.vsg_z[i], c.w.pvt_z[i]);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_vec_tend(const CtxT<R> c, const StateDev old, const StateDev nw, int rk) {
  b_vec_tend(c, old, nw, rk, bid_here());
}

// The first substep's pressure gradient and wind update in one kernel: the pressure-gradient terms
// of a cell's u, v and w points are stored for the second substep and used from registers for the
// first one.  The thread of column 1 also updates the u point of column nx+1, and the thread of
// row 1 the v point of row ny+1, whose pressure-gradient terms are the copies of theirs.
template <class R>
__device__ __forceinline__ void b_pgrad_vt(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int lfirst, const Bid bid) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  const int nx = c.G.nx, ny = c.G.ny;
  if (jk > nx || ji > ny || jl > c.G.nl) return;
  const PGrad<R> g = pgrad_interior(c, old, lfirst, ji, jk, jl);
  const VtWeights<R> w = vt_weights(c, 1);
  const ix_t i = LG_IX(c.G, ji, jk, jl);
  vt_u(c, old, nw, w, i, g.old_u, g.nl_u, g.l_u, c.w.vsg_x[i], c.w.pvt_x[i]);
  if (jk == 1) {
    const ix_t iX = LG_IX(c.G, ji, nx + 1, jl);
    vt_u(c, old, nw, w, iX, g.old_u, g.nl_u, g.l_u, c.w.vsg_x[iX], c.w.pvt_x[iX]);
  }
  vt_v(c, old, nw, w, i, g.old_v, g.nl_v, g.l_v, c.w.vsg_y[i], c.w.pvt_y[i]);
  if (ji == 1) {
    const ix_t iY = LG_IX(c.G, ny + 1, jk, jl);
    vt_v(c, old, nw, w, iY, g.old_v, g.nl_v, g.l_v, c.w.vsg_y[iY], c.w.pvt_y[iY]);
  }
  if (jl >= 2) vt_w(c, w, ji, jk, jl, i, g.nl_w, g.l_w, c.w.vsg_z[i], c.w.pvt_z[i]);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_pgrad_vt(const CtxT<R> c, const StateDev old, const StateDev nw, int lfirst) {
  b_pgrad_vt(c, old, nw, lfirst, bid_here());
}

// The second substep's vorticity flux term (with the kinetic-energy gradient) and wind update in
// one kernel: both terms are stored for the next step's first substep and used from registers
// here.  The thread of column 1 also updates the u point of column nx+1, and the thread of row 1
// the v point of row ny+1.  Without the kinetic-energy gradient (linear runs) its stored values
// are read.
template <class R>
__device__ __forceinline__ void b_vort_flux_vt(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int with_ke_grad,
                                               const Bid bid) {
  KIDX_ZFAST(c.G.nx, c.G.ny, c.G.nl);
  const int nx = c.G.nx, ny = c.G.ny;
  if (jk > nx || ji > ny || jl > c.G.nl) return;
  const ix_t i = LG_IX(c.G, ji, jk, jl), iX = LG_IX(c.G, ji, nx + 1, jl), iY = LG_IX(c.G, ny + 1, jk, jl);
  const VortFlux<R> f = vort_flux_interior(c, ji, jk, jl);
  KeGrad<R> g;
  if (with_ke_grad) {
    g = ke_grad_interior(c, ji, jk, jl);
  } else {
    g.x = c.w.vsg_x[i];
    g.y = c.w.vsg_y[i];
    g.z = jl >= 2 ? c.w.vsg_z[i] : R(0.0);
    g.x_edge = jk == 1 ? c.w.vsg_x[iX] : R(0.0);
    g.y_edge = ji == 1 ? c.w.vsg_y[iY] : R(0.0);
  }
  const VtWeights<R> w = vt_weights(c, 2);
  vt_u(c, old, nw, w, i, c.w.pold_u[i], c.w.pnl_u[i], c.w.pl_u[i], g.x, f.x);
  if (jk == 1) vt_u(c, old, nw, w, iX, c.w.pold_u[iX], c.w.pnl_u[iX], c.w.pl_u[iX], g.x_edge, f.x);
  vt_v(c, old, nw, w, i, c.w.pold_v[i], c.w.pnl_v[i], c.w.pl_v[i], g.y, f.y);
  if (ji == 1) vt_v(c, old, nw, w, iY, c.w.pold_v[iY], c.w.pnl_v[iY], c.w.pl_v[iY], g.y_edge, f.y);
  if (jl >= 2) vt_w(c, w, ji, jk, jl, i, c.w.pnl_w[i], c.w.pl_w[i], g.z, f.z);
}
template <class R>
__global__ void __launch_bounds__(BX * BY) k_vort_flux_vt(const CtxT<R> c, const StateDev old, const StateDev nw, int with_ke_grad) {
  b_vort_flux_vt(c, old, nw, with_ke_grad, bid_here());
}

// scalar_tend_expl (mo_scalar_tend_expl.F90:24-181) is the column-marching k_scalar_col of
// lgame_scalar.hip.

// The plain form (FP64 stencils) and the mixed form (FP32 stencils).
#define LG_INSTANTIATE(R)                                                                                  \
  template __global__ void k_temperature<R>(const CtxT<R>, const StateDev);                              \
  template __global__ void k_theta_grad<R>(const CtxT<R>, const StateDev, int);                          \
  template __global__ void k_theta_div<R>(const CtxT<R>, int);                                           \
  template __global__ void k_theta_edge<R>(const CtxT<R>, const StateDev, int);                          \
  template __global__ void k_pot_vort<R>(const CtxT<R>, const StateDev, int);                            \
  template __global__ void k_pot_vort2<R>(const CtxT<R>, const StateDev, int);                           \
  template __global__ void k_vort_flux<R>(const CtxT<R>, int);                                           \
  template __global__ void k_dissipation<R>(const CtxT<R>, const StateDev);                              \
  template __global__ void k_vec_tend<R>(const CtxT<R>, const StateDev, const StateDev, int);                 \
  template __global__ void k_pgrad_vt<R>(const CtxT<R>, const StateDev, const StateDev, int);                 \
  template __global__ void k_vort_flux_vt<R>(const CtxT<R>, const StateDev, const StateDev, int);

LG_INSTANTIATE(double)
LG_INSTANTIATE(float)
#undef LG_INSTANTIATE

// ---- the same kernels over a group of domains (one launch per phase) ------------------------
LG_BATCH_KERNEL(temperature, BX * BY, PH_TEMPERATURE, b_temperature(d.c, d.s[a.old_slot], bid))
LG_BATCH_KERNEL(theta_grad, BX * BY, PH_THETA_GRAD, b_theta_grad(d.c, d.s[a.state_slot], a.nfield, bid))
LG_BATCH_KERNEL(theta_div, BX * BY, PH_THETA_DIV, b_theta_div(d.c, a.nfield, bid))
LG_BATCH_KERNEL(theta_edge, BX * BY, PH_THETA_EDGE, b_theta_edge(d.c, d.s[a.state_slot], a.nfield, bid))
LG_BATCH_KERNEL(pgrad, BX * BY, PH_PGRAD, b_pgrad_vt(d.c, d.s[a.old_slot], d.s[a.new_slot], a.lfirst, bid))
LG_BATCH_KERNEL(pot_vort, BX * BY, PH_POT_VORT, b_pot_vort(d.c, d.s[a.state_slot], a.ke, bid))
LG_BATCH_KERNEL(vort_flux, BX * BY, PH_VORT_FLUX, b_vort_flux(d.c, a.ke, bid))
LG_BATCH_KERNEL(vort_flux_vt, BX * BY, PH_VORT_FLUX_VT, b_vort_flux_vt(d.c, d.s[a.old_slot], d.s[a.new_slot], a.ke, bid))
LG_BATCH_KERNEL(dissipation, BX * BY, PH_DISSIPATION, b_dissipation(d.c, d.s[a.old_slot], bid))
LG_BATCH_KERNEL(vec_tend, BX * BY, PH_VEC_TEND, b_vec_tend(d.c, d.s[a.old_slot], d.s[a.new_slot], a.rk, bid))

template <class R>
bool batch_launch_step(int ph, dim3 g, dim3 b, hipStream_t st, const MemberDev<R> *M, const int *prefix, int nm, const BatchArgs &a) {
#define LG_B(P, K)                                                  \
  case P:                                                           \
    hipLaunchKernelGGL(K<R>, g, b, 0, st, M, prefix, nm, a);        \
    return true;
  switch (ph) {
    LG_B(PH_TEMPERATURE, kb_temperature)
    LG_B(PH_THETA_GRAD, kb_theta_grad)
    LG_B(PH_THETA_DIV, kb_theta_div)
    LG_B(PH_THETA_EDGE, kb_theta_edge)
    LG_B(PH_PGRAD, kb_pgrad)
    LG_B(PH_POT_VORT, kb_pot_vort)
    LG_B(PH_VORT_FLUX, kb_vort_flux)
    LG_B(PH_VORT_FLUX_VT, kb_vort_flux_vt)
    LG_B(PH_DISSIPATION, kb_dissipation)
    LG_B(PH_VEC_TEND, kb_vec_tend)
    default: return false;
  }
#undef LG_B
}
template bool batch_launch_step<double>(int, dim3, dim3, hipStream_t, const MemberDev<double> *, const int *, int, const BatchArgs &);
template bool batch_launch_step<float>(int, dim3, dim3, hipStream_t, const MemberDev<float> *, const int *, int, const BatchArgs &);

}  // namespace lgame
