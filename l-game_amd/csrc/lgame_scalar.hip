// lgame_scalar.hip -- scalar_tend_expl as one column-marching kernel (gfx950).
//
// scalar_tend_expl (mo_scalar_tend_expl.F90:24-181) forms the gas and theta_v mass fluxes at the
// u and v faces (scalar_times_vector_h, mo_multiplications.F90), their divergence with the
// terrain-following contravariant corrections (div_h, mo_divergence_operators.F90:16-68, with
// vertical_contravariant_corr, mo_averaging.F90:16-54) and the upstream tracer divergences
// (div_h_tracers, :70-147).  Done one level per thread, every face flux is formed by two cells,
// every interface correction by the cells above and below it, and each layer's faces are loaded
// by three levels.  Here one lane owns one column and marches down it:
// - the lanes of a 64-lane row hold 64 consecutive columns; 62 of them are cells, the first and
//   the last supply the west and east halo.  A lane forms the fluxes of its own west face; the
//   east face of a cell is the west face of the next lane (a cross-lane shift, no memory).  The
//   north and south neighbours are loads of the rows above and below;
// - each layer is prepared once, one level ahead of the cell being finished: its faces, and the
//   contravariant correction of the interface above it, formed from that layer and the one
//   above, which both adjacent cells then use;
// - the tracer densities of the layers above and below stay in registers for the upstream
//   choice at the interfaces.
// The expressions and their order are those of the per-point kernels (and of the reference), so
// the results are the same bits.  The column can be cut into segments (each starts with one
// extra layer) so that small domains still fill the GPU.
#include "lgame_ops.h"
#include "lgame_step.h"
#include "lgame_batch.h"

namespace lgame {

namespace {

// the value of `x` at the next (down) or previous (up) lane of the wavefront
template <class T>
__device__ __forceinline__ T from_next_lane(T x) {
  return __shfl_down(x, 1, 64);
}
template <class T>
__device__ __forceinline__ T from_prev_lane(T x) {
  return __shfl_up(x, 1, 64);
}

// the constituents: the moist layout is fixed at 5 condensates, the gas and the vapour (lgame_init
// accepts no other), the dry one is the gas alone
template <bool TR>
struct Cons {
  static constexpr int nc = TR ? 7 : 1, ncond = TR ? 5 : 0;
  static constexpr bool tracer(int q) { return TR && q != ncond; }
};

// this lane's place: row ji, lane column jk (0..nx+1 and beyond), the state column js (jk
// wrapped), the metric column jkm (jk clamped to the stored u points), the wrapped rows jmr
// (ji-1) and j1 (ji+1), and whether the lane's column is a cell of this row
struct Lane {
  int ji, jk, js, jkm, jmr, j1;
  bool cell;
};

// the correction of one interface for the three flux kinds
template <class R>
struct Corr {
  R g, h, w;
};

// one prepared layer of this lane's column
template <class R, bool TR>
struct Prep {
  R hg, hh;                 // the horizontal part (comp_h) of the gas and theta_v divergences
  R ht[TR ? 7 : 1];         // the same for the tracers (upstream face densities)
  R r[TR ? 7 : 1];          // own densities of the constituents other than the gas
  R rg;                     // own gas density
  Corr<R> P;                // vertical_contravariant_corr terms of this layer alone
  Corr<R> I;                // correction of the interface at the top of this layer
  R wfs;                    // w_free_slip of this layer's winds (used at the lowest layer)
};

// Faces, fluxes and the horizontal divergence parts of layer l; Pa holds the correction terms
// of the layer above (l-1), from which the correction of interface l is formed.  Every load of
// the layer is made unconditionally (indices are valid at levels 1..n_layers+1) so that they
// issue together.
template <class R, bool TR, bool O3>
__device__ __forceinline__ void prep_layer(const CtxT<R> &c, const StateDev &ss, const StateDev &nw, const Lane &p, int l,
                                           const Corr<R> &Pa, Prep<R, TR> &o) {
  using K = Cons<TR>;
  const Geom &G = c.G;
  const Fld<const double> rg = ss.rho[K::ncond], thp = ss.theta_v_pert;
  const ix_t iC = LG_IX(G, p.ji, p.js, l), iN = LG_IX(G, p.jmr, p.js, l), iS = LG_IX(G, p.j1, p.js, l);
  const ix_t iM = LG_IX(G, p.ji, p.jkm, l), iMs = LG_IX(G, p.ji + 1, p.jkm, l);
  // west face (u at column jk; the column nx+1 lane forms the copy of face 1)
  const R rC = R(rg[iC]);
  const R thC = c.g.theta_v_bg[iC] + R(thp[iC]);
  const R rW = from_prev_lane(rC), thW = from_prev_lane(thC);
  const R uS = R(nw.wind_u[iC]);
  const R fu = R(0.5) * (rW + rC) * uS;
  const R thu = O3 ? c.w.tu[0][iC] * fu : R(0.5) * (thW + thC) * fu;
  // north face (v at row ji)
  const R rN = R(rg[iN]);
  const R thN = c.g.theta_v_bg[iN] + R(thp[iN]);
  const R vN = R(nw.wind_v[iC]);
  const R fv = R(0.5) * (rN + rC) * vN;
  const R thv = O3 ? c.w.tv[0][iC] * fv : R(0.5) * (thN + thC) * fv;
  // south face (v at row ji+1, the copy of row 1 at ji = ny)
  const R rS = R(rg[iS]);
  const R thS = c.g.theta_v_bg[iS] + R(thp[iS]);
  const R vSw = R(nw.wind_v[iS]);
  const R fvs = R(0.5) * (rC + rS) * vSw;
  const R thvs = O3 ? c.w.tv[0][iS] * fvs : R(0.5) * (thC + thS) * fvs;
  // the winds as stored at this lane's u point and at row ji+1 (they differ from the wrapped ones
  // only at column nx+1 and row ny+1)
  const R uM = R(nw.wind_u[iM]);
  const R vSm = R(nw.wind_v[LG_IX(G, p.ji + 1, p.js, l)]);
  const Faces<R> fg{from_next_lane(fu), fv, fu, fvs};
  const Faces<R> fh{from_next_lane(thu), thv, thu, thvs};
  const Faces<R> fw{from_next_lane(uM), vN, uM, vSm};
  const R ue = from_next_lane(uS);
  // metrics of the four faces
  const R ax0 = c.g.area_x[iM], ay0 = c.g.area_y[iM], ay1 = c.g.area_y[iMs];
  const R ax1 = from_next_lane(ax0);
  const R sx0 = c.g.slope_x[iM];
  const VccGeo<R> geo{from_next_lane(sx0), c.g.slope_y[iM], sx0, c.g.slope_y[iMs],
                      c.g.ipw[0][iM],       c.g.ipw[1][iM],   c.g.ipw[2][iM], c.g.ipw[3][iM]};
  // div_h's horizontal part and the contravariant corrections
  o.hg = fg.e * ax1 + fg.n * ay0 - fg.w * ax0 - fg.s * ay1;
  o.hh = fh.e * ax1 + fh.n * ay0 - fh.w * ax0 - fh.s * ay1;
  const Corr<R> zero{R(0.0), R(0.0), R(0.0)};
  const Corr<R> &a = l != G.nflat + 1 ? Pa : zero;
  o.I = Corr<R>{vcc_sub(a.g, geo, fg), vcc_sub(a.h, geo, fh), TR ? vcc_sub(a.w, geo, fw) : R(0.0)};
  o.P = Corr<R>{vcc_sub(R(0.0), geo, fg), vcc_sub(R(0.0), geo, fh), TR ? vcc_sub(R(0.0), geo, fw) : R(0.0)};
  o.wfs = geo.w0 * fw.e * geo.sx1 + geo.w1 * fw.n * geo.sy0 + geo.w2 * fw.w * geo.sx0 + geo.w3 * fw.s * geo.sy1;
  o.rg = rC;
  // the tracers' upstream face densities (div_h_tracers, mo_divergence_operators.F90:70-147)
#pragma unroll
  for (int q = 0; q < K::nc; ++q) {
    if (q == K::ncond) continue;
    o.r[q] = R(ss.rho[q][iC]);
  }
  if (TR) {
    const R uw = fw.w, vn = fw.n, vs = vSw;
#pragma unroll
    for (int q = 0; q < K::nc; ++q) {
      if (!K::tracer(q)) continue;
      const R r0 = o.r[q], rE = from_next_lane(r0), rWq = from_prev_lane(r0);
      const R rNq = R(ss.rho[q][iN]), rSq = R(ss.rho[q][iS]);
      const R v_e = ue >= R(0.0) ? r0 : rE;
      const R v_w = uw >= R(0.0) ? rWq : r0;
      const R v_n = vn >= R(0.0) ? r0 : rNq;
      const R v_s = vs >= R(0.0) ? rSq : r0;
      o.ht[q] = (v_e * ue) * ax1 + (v_n * vn) * ay0 - (v_w * uw) * ax0 - (v_s * vs) * ay1;
    }
  }
}

// c_v_mass_weighted_air (mo_derived.F90:157-187) from the densities of the cell
template <class R, bool TR>
__device__ __forceinline__ R cv_air(const Prep<R, TR> &L, R temperature) {
  using K = Cons<TR>;
  R r = R(0.0);
  if (TR) {
#pragma unroll
    for (int jc = 1; jc <= K::ncond; ++jc) r = r + L.r[jc - 1] * c_p_cond(jc, temperature);
    r = r + (L.rg - L.r[K::ncond + 1]) * R(C_D_V);
    r = r + L.r[K::ncond + 1] * R(C_V_V);
  } else {
    r = L.rg * R(C_D_V);
  }
  return r;
}

// the cell's own inputs of finish_cell, loaded before the next layer is prepared so that their
// latency is hidden behind that work, and before any store of the cell (the tendency arrays
// are read and written)
template <class R, bool TR>
struct CellIn {
  R az_u, az_l, V, heat, temperature, exner, md_g, theta;
  R told[TR ? 7 : 1], md[TR ? 7 : 1];
};

template <class R, bool TR, bool DIFF>
__device__ __forceinline__ void load_cell(const CtxT<R> &c, const StateDev &ss, const Lane &p, int jl, CellIn<R, TR> &f) {
  using K = Cons<TR>;
  const Geom &G = c.G;
  const ix_t i = LG_IX(G, p.ji, p.js, jl);
  f.az_u = c.g.area_z[i];
  f.az_l = c.g.area_z[i + (ix_t)G.lvl];
  f.V = c.g.volume[i];
  f.heat = c.w.heating_diss[i] + c.w.radiation_tendency[i] + c.w.csh[i];
  if (DIFF && c.C.ltemp_h) f.heat = f.heat + c.w.tdh[i];  // diag%temp_diff_heating (of this step's first substep)
  f.temperature = c.w.temperature[i];
  f.exner = c.g.exner_bg[i] + R(ss.exner_pert[i]);
  f.md_g = R(0.0);
  f.theta = R(0.0);
  if (DIFF && c.C.lmass_h) {
    f.md_g = c.w.mdt[K::ncond][i];
    f.theta = c.g.theta_v_bg[i] + R(ss.theta_v_pert[i]);
  }
#pragma unroll
  for (int q = 0; q < K::nc; ++q) {
    if (!K::tracer(q)) continue;
    f.told[q] = c.w.tend_rho[q][i];
    f.md[q] = DIFF && c.C.lmass_h ? c.w.mdt[q][i] : R(0.0);
  }
}

// the tendencies of cell (ji, jk, jl) from its prepared layer L, its inputs f, the corrections of
// its upper (L.I) and lower (Ilo) interface and the tracer densities of the layers above and below
template <class R, bool TR, bool DIFF>
__device__ __forceinline__ void finish_cell(const CtxT<R> &c, const StateDev &nw, int rk, const Lane &p, int jl,
                                            const Prep<R, TR> &L, const CellIn<R, TR> &f, const Corr<R> &Ilo, const R *r_up,
                                            const R *r_lo) {
  using K = Cons<TR>;
  const Geom &G = c.G;
  const int nl = G.nl, nflat = G.nflat, ncond = K::ncond;
  const bool up = jl > nflat, lo = jl >= nflat && jl < nl;
  const ix_t i = LG_IX(G, p.ji, p.js, jl);
  const R az_u = up ? f.az_u : R(0.0);
  const R az_l = lo ? f.az_l : R(0.0);
  const R V = f.V;
  const R cu_g = up ? L.I.g : R(0.0), cu_h = up ? L.I.h : R(0.0), cu_w = up ? L.I.w : R(0.0);
  const R cl_g = lo ? Ilo.g : R(0.0), cl_h = lo ? Ilo.h : R(0.0), cl_w = lo ? Ilo.w : R(0.0);
  auto div = [&](R comp_h, R cu, R cl) {
    R comp_v = R(0.0);
    if (jl == nflat) {
      comp_v = -cl * az_l;
    } else if (jl == nl) {
      comp_v = cu * az_u;
    } else if (jl > nflat) {
      comp_v = cu * az_u - cl * az_l;
    }
    return (comp_h + comp_v) / V;
  };
  const R div_g = div(L.hg, cu_g, cl_g);
  const R div_h_ = div(L.hh, cu_h, cl_h);
  R rt_tend = -div_h_ + R(C_D_V) * L.rg / cv_air(L, f.temperature) * f.heat / (R(C_D_P) * f.exner);
  R tg = -div_g;
  if (DIFF && c.C.lmass_h) {  // diag%mass_diff_tendency of the gas, times theta_v in the rhotheta_v tendency
    tg = tg + f.md_g;
    rt_tend = rt_tend + (R(0.0) + f.md_g) * f.theta;
  }
  R tt[TR ? 7 : 1];
  if (TR) {
    const R new_w = rk == 2 ? R(0.5) : R(1.0), old_w = R(1.0) - new_w;
    const R cu = (jl == nl || jl > nflat) && jl != nflat ? cu_w : R(0.0);
    const R cl = jl == nflat || (jl > nflat && jl != nl) ? cl_w : R(0.0);
#pragma unroll
    for (int q = 0; q < K::nc; ++q) {
      if (!K::tracer(q)) continue;
      const R rC = L.r[q];
      const R v_u = cu <= R(0.0) ? r_up[q] : rC;
      const R v_l = cl <= R(0.0) ? rC : r_lo[q];
      R comp_v = R(0.0);
      if (jl == nflat) {
        comp_v = -v_l * cl * az_l;
      } else if (jl == nl) {
        comp_v = v_u * cu * az_u;
      } else if (jl > nflat) {
        comp_v = v_u * cu * az_u - v_l * cl * az_l;
      }
      const R d = (L.ht[q] + comp_v) / V;
      const R x = DIFF && c.C.lmass_h ? -d + f.md[q] : -d;
      tt[q] = old_w * f.told[q] + new_w * x;
    }
  }
  if (!p.cell) return;
  c.w.tend_rho[ncond][i] = tg;
  c.w.tend_rhotheta_v[i] = rt_tend;
#pragma unroll
  for (int q = 0; q < K::nc; ++q)
    if (K::tracer(q)) c.w.tend_rho[q][i] = tt[q];
  if (c.C.lfreeslip && jl == nl) nw.wind_w[LG_IX(G, p.ji, p.js, nl + 1)] = L.wfs;  // w_free_slip, mo_inner_product.F90:52-75
}

}  // namespace

// Launch: a 1-D grid of ceil(nx/62) x ceil(ny/by) x nseg blocks of 64 x by threads (by rows of
// 64 lanes), dealt to the XCDs so that each takes a contiguous range of tiles with x fastest:
// the rows above and below a tile are read by the blocks next to it on the same XCD.
template <class R, bool TR, bool O3, bool DIFF>
__device__ __forceinline__ void b_scalar_col(const CtxT<R> &c, const StateDev &ss, const StateDev &nw, int rk, int nseg, const Bid bid) {
  const Geom &G = c.G;
  const int nx = G.nx, ny = G.ny, nl = G.nl;
  const int tx_n = (nx + SCOL_CELLS - 1) / SCOL_CELLS, ty_n = (ny + (int)blockDim.y - 1) / (int)blockDim.y;
  const int n = tx_n * ty_n * nseg, b = bid.x;
  const int k = b % N_XCD, s = b / N_XCD, q = n / N_XCD, rr = n % N_XCD;
  const int t = k * q + (k < rr ? k : rr) + s;
  const int tx = t % tx_n, ty = (t / tx_n) % ty_n, seg = t / (tx_n * ty_n);
  Lane p;
  p.ji = ty * (int)blockDim.y + (int)threadIdx.y + 1;
  if (p.ji > ny) return;  // a whole row (wavefront) at once
  const int lane = (int)threadIdx.x;
  p.jk = tx * SCOL_CELLS + lane;  // lane 0: the west halo column
  p.js = p.jk < 1 ? nx : (p.jk > nx ? 1 : p.jk);
  p.jkm = p.jk < 1 ? 1 : (p.jk > nx + 1 ? nx + 1 : p.jk);
  p.jmr = wrapm(p.ji, ny);
  p.j1 = p.ji == ny ? 1 : p.ji + 1;
  p.cell = lane >= 1 && lane <= SCOL_CELLS && p.jk <= nx;
  // this segment's levels
  const int l0 = 1 + (nl * seg) / nseg, l1 = (nl * (seg + 1)) / nseg;
  const Corr<R> zero{R(0.0), R(0.0), R(0.0)};
  Prep<R, TR> cur, nxt;
  R r_up[TR ? 7 : 1];
  for (R &x : r_up) x = R(0.0);
  // prologue: the layer above the segment (its correction terms and tracer densities), then the
  // segment's first layer
  Corr<R> Pa = zero;
  if (l0 > 1) {
    prep_layer<R, TR, O3>(c, ss, nw, p, l0 - 1, zero, nxt);
    Pa = nxt.P;
#pragma unroll
    for (int qq = 0; qq < (TR ? 7 : 1); ++qq) r_up[qq] = nxt.r[qq];
  }
  prep_layer<R, TR, O3>(c, ss, nw, p, l0, Pa, cur);
  for (int jl = l0; jl <= l1; ++jl) {
    // the next layer (at jl = n_layers the padding level n_layers+1, whose values are not used)
    CellIn<R, TR> f;
    load_cell<R, TR, DIFF>(c, ss, p, jl, f);
    prep_layer<R, TR, O3>(c, ss, nw, p, jl + 1, cur.P, nxt);
    finish_cell<R, TR, DIFF>(c, nw, rk, p, jl, cur, f, nxt.I, r_up, nxt.r);
#pragma unroll
    for (int qq = 0; qq < (TR ? 7 : 1); ++qq) r_up[qq] = cur.r[qq];
    cur = nxt;
  }
}

// DIFF: the step has temperature or mass diffusion, whose inputs the cells read.  Without them the
// FP64 moist kernel needs 272 registers instead of 290 and runs 7% faster; the FP32 one was no
// faster (1% slower, also when held to three wavefronts per SIMD, which then fits), so it always
// runs the variant that checks the switches at run time.
template <class R>
__host__ __device__ constexpr bool scol_diff_variant(bool diff) {
  return diff || sizeof(R) == 4;
}
template <class R, bool TR, bool O3, bool DIFF>
__global__ void __launch_bounds__(BX * 4)
    k_scalar_col(const CtxT<R> c, const StateDev ss, const StateDev nw, int rk, int nseg) {
  b_scalar_col<R, TR, O3, DIFF>(c, ss, nw, rk, nseg, bid_here());
}

template <class R>
void launch_scalar_col(const CtxT<R> &c, const StateDev &ss, const StateDev &nw, int rk, int rows, int nseg, hipStream_t st) {
  const int tx_n = (c.G.nx + SCOL_CELLS - 1) / SCOL_CELLS, ty_n = (c.G.ny + rows - 1) / rows;
  const dim3 blk(BX, rows, 1), grd((unsigned)(tx_n * ty_n * nseg));
  const bool tr = c.G.nc > 1, o3 = c.C.order == 3, diff = scol_diff_variant<R>(c.C.ltemp_h || c.C.lmass_h);
#define LG_SCOL(TR, O3, DIFF) hipLaunchKernelGGL((k_scalar_col<R, TR, O3, DIFF>), grd, blk, 0, st, c, ss, nw, rk, nseg)
  if (diff) {
    if (tr && o3) LG_SCOL(true, true, true);
    else if (tr) LG_SCOL(true, false, true);
    else if (o3) LG_SCOL(false, true, true);
    else LG_SCOL(false, false, true);
  } else {
    if (tr && o3) LG_SCOL(true, true, false);
    else if (tr) LG_SCOL(true, false, false);
    else if (o3) LG_SCOL(false, true, false);
    else LG_SCOL(false, false, false);
  }
#undef LG_SCOL
}

template void launch_scalar_col<double>(const CtxT<double> &, const StateDev &, const StateDev &, int, int, int, hipStream_t);
template void launch_scalar_col<float>(const CtxT<float> &, const StateDev &, const StateDev &, int, int, int, hipStream_t);

// ---- the same kernel over a group of domains (one launch) --------------------------------------
template <class R, bool TR, bool O3, bool DIFF>
__global__ void __launch_bounds__(BX * 4)
    kb_scalar_col(const MemberDev<R> *__restrict__ M, const int *__restrict__ prefix, int nm, const BatchArgs a) {
  Bid bid;
  const MemberDev<R> &d = batch_block(M, prefix, nm, PH_SCALAR, bid);
  b_scalar_col<R, TR, O3, DIFF>(d.c, d.s[a.state_slot], d.s[a.new_slot], a.rk, d.seg, bid);
}

template <class R>
void batch_launch_scalar(dim3 g, dim3 b, hipStream_t st, const MemberDev<R> *M, const int *prefix, int nm, const BatchArgs &a,
                         bool tr, bool o3, bool diff) {
  diff = scol_diff_variant<R>(diff);
#define LG_SCOL(TR, O3, DIFF) hipLaunchKernelGGL((kb_scalar_col<R, TR, O3, DIFF>), g, b, 0, st, M, prefix, nm, a)
  if (diff) {
    if (tr && o3) LG_SCOL(true, true, true);
    else if (tr) LG_SCOL(true, false, true);
    else if (o3) LG_SCOL(false, true, true);
    else LG_SCOL(false, false, true);
  } else {
    if (tr && o3) LG_SCOL(true, true, false);
    else if (tr) LG_SCOL(true, false, false);
    else if (o3) LG_SCOL(false, true, false);
    else LG_SCOL(false, false, false);
  }
#undef LG_SCOL
}
template void batch_launch_scalar<double>(dim3, dim3, hipStream_t, const MemberDev<double> *, const int *, int, const BatchArgs &,
                                          bool, bool, bool);
template void batch_launch_scalar<float>(dim3, dim3, hipStream_t, const MemberDev<float> *, const int *, int, const BatchArgs &,
                                         bool, bool, bool);

}  // namespace lgame
