// lgame_batch.h -- one launch per kernel phase for a group of domains (an ensemble).
//
// An ensemble's members whose steps run the same kernel sequence form a group.  For each phase of
// the step (one kernel of enqueue_pchevi) the group launches one 1-D grid that holds every
// member's blocks back to back: `prefix` gives the first block of each member, a block finds its
// member by a binary search over it (the table is wave-uniform, so the search runs on the scalar
// unit), then its coordinates inside that member's own launch of the phase, and runs the same
// kernel body as the single-domain launch.  The members' contexts and state slots live in a
// device array (MemberDev) that the blocks read through the scalar cache.
#pragma once

#include "lgame_ops.h"
#include "lgame_rad.h"
#include "lgame_step.h"

namespace lgame {

// PH_PGRAD is the first substep's pressure gradient with its wind update (b_pgrad_vt), PH_VORT_FLUX_VT
// the second substep's vorticity flux term with its wind update (b_vort_flux_vt); PH_VEC_TEND is the
// second substep's wind update alone (runs without momentum advection).
enum Phase : int {
  PH_TEMPERATURE, PH_SFC_TURB, PH_THETA_GRAD, PH_THETA_DIV, PH_THETA_EDGE, PH_PGRAD, PH_POT_VORT,
  PH_VORT_FLUX, PH_DIFF_ZETA, PH_DIFF_COEF, PH_DIFF_DUAL, PH_DIFF_MOM, PH_PBL_WIND,
  PH_DISSIPATION, PH_VEC_TEND, PH_DIFF_FLUX, PH_DIFF_DIV, PH_SCALAR, PH_SURFACE, PH_SOLVE_VER, PH_SOLVE_GEN, PH_RAD,
  PH_VORT_FLUX_VT,
  PH_COUNT
};

// the tile-table entry that gives a phase its rows per block (-1: a column kernel)
inline int phase_tile(int ph) {
  switch (ph) {
    case PH_TEMPERATURE: return KI_TEMPERATURE;
    case PH_SFC_TURB: case PH_PBL_WIND: case PH_SURFACE: case PH_DIFF_DUAL: return KI_DIFF_DUAL;
    case PH_THETA_GRAD: return KI_THETA_GRAD;
    case PH_THETA_DIV: return KI_THETA_DIV;
    case PH_THETA_EDGE: return KI_THETA_EDGE;
    case PH_PGRAD: return KI_PGRAD;
    case PH_POT_VORT: return KI_POT_VORT;
    case PH_VORT_FLUX: case PH_VORT_FLUX_VT: return KI_VORT_FLUX;
    case PH_DIFF_ZETA: return KI_DIFF_ZETA;
    case PH_DIFF_COEF: return KI_DIFF_COEF;
    case PH_DIFF_MOM: return KI_DIFF_MOM;
    case PH_DISSIPATION: return KI_DISSIPATION;
    case PH_VEC_TEND: return KI_VEC_TEND;
    case PH_DIFF_FLUX: return KI_DIFF_FLUX;
    case PH_DIFF_DIV: return KI_DIFF_DIV;
    case PH_SCALAR: return KI_SCALAR_DIV;
    default: return -1;
  }
}

// One member on the device: its context, its two state slots, its radiation data (pointers
// into the group's surrogate buffers) and the time of the current radiation call.
template <class R>
struct MemberDev {
  CtxT<R> c;
  StateDev s[2];
  RadDev rad;
  double t_rad;
  int seg;  // vertical segments of the column-marching scalar kernel
};

// the arguments that are the same for every member of a launch
struct BatchArgs {
  int old_slot, new_slot;  // state_old and state_new
  int state_slot;          // the slot of the phase's state argument (state_old, or state_new at rk 2)
  int rk, lfirst, nfield, flags;
  int ke;  // the mass-flux kernel also forms the kinetic energy, the vorticity-flux kernel its gradient
};

// Blocks of phase ph for one member as (gx, gy, gz), with `by` rows per block for the stencil
// phases; phases launched on a 1-D grid have gy = gz = 1.  Mirrors the launches of
// enqueue_pchevi.
__host__ __device__ inline void phase_grid(int ph, const Geom &G, int by, int seg, int &gx, int &gy, int &gz) {
  const int nx = G.nx, ny = G.ny, nl = G.nl, nlev = G.nlev;
  int ex = nx, ey = ny, ez = nl;
  bool flat = false;  // a 1-D grid (tile_zfast kernels)
  gx = gy = gz = 0;
  switch (ph) {
    case PH_TEMPERATURE: case PH_THETA_DIV: case PH_DIFF_DUAL: case PH_DISSIPATION: break;
    case PH_SFC_TURB: case PH_PBL_WIND: case PH_SURFACE: ez = 1; break;
    case PH_THETA_GRAD: case PH_THETA_EDGE: case PH_VEC_TEND: ex = nx + 1; ey = ny + 1; break;
    case PH_PGRAD: case PH_VORT_FLUX: case PH_VORT_FLUX_VT: case PH_DIFF_COEF: case PH_DIFF_MOM: case PH_DIFF_FLUX: case PH_DIFF_DIV: flat = true; break;
    case PH_POT_VORT: ez = nlev; flat = true; break;
    case PH_DIFF_ZETA: ex = nx + 1; ey = ny + 1; flat = true; break;
    case PH_SCALAR: gx = ((nx + SCOL_CELLS - 1) / SCOL_CELLS) * ((ny + by - 1) / by) * seg; gy = gz = 1; return;
    case PH_SOLVE_VER: gx = col_waves(G); gy = gz = 1; return;
    case PH_RAD: gx = (nx * ny + BCOL - 1) / BCOL; gy = gz = 1; return;
    case PH_SOLVE_GEN: gx = (col_waves(G) + GEN_GROUPS - 1) / GEN_GROUPS; gy = gz = 1; return;
    default: return;
  }
  gx = (ex + BX - 1) / BX;
  gy = (ey + by - 1) / by;
  gz = ez;
  if (flat) {
    gx = gx * gy * gz;
    gy = gz = 1;
  }
}

// the member of block b: prefix[m] <= b < prefix[m+1]
__device__ __forceinline__ int batch_member(const int *__restrict__ prefix, int nm, int b) {
  int lo = 0, hi = nm;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// this block's member and its coordinates inside the member's own launch of phase ph
template <class R>
__device__ __forceinline__ const MemberDev<R> &batch_block(const MemberDev<R> *__restrict__ M, const int *__restrict__ prefix,
                                                           int nm, int ph, Bid &bid) {
  const int b = (int)blockIdx.x, m = batch_member(prefix, nm, b);
  const MemberDev<R> &d = M[m];
  int gx, gy, gz;
  phase_grid(ph, d.c.G, (int)blockDim.y, d.seg, gx, gy, gz);
  const int local = b - prefix[m];
  bid.x = local % gx;
  bid.y = (local / gx) % gy;
  bid.z = local / (gx * gy);
  return d;
}

// A batched kernel: every block runs `CALL` for its member d with its coordinates bid.
#define LG_BATCH_KERNEL(NAME, BOUNDS, PH, CALL)                                                                        \
  template <class R>                                                                                                  \
  __global__ void __launch_bounds__(BOUNDS)                                                                           \
      kb_##NAME(const MemberDev<R> *__restrict__ M, const int *__restrict__ prefix, int nm, const BatchArgs a) {      \
    Bid bid;                                                                                                          \
    const MemberDev<R> &d = batch_block(M, prefix, nm, PH, bid);                                                      \
    CALL;                                                                                                             \
  }

// Launch phase ph of a group on `st`: grid g (the group's total blocks), block b.  Each source
// file launches the phases whose kernels it holds and returns false for the others.
#define LG_BATCH_LAUNCHER(NAME)                                                                                      \
  template <class R>                                                                                                  \
  bool NAME(int ph, dim3 g, dim3 b, hipStream_t st, const MemberDev<R> *M, const int *prefix, int nm, const BatchArgs &a);
LG_BATCH_LAUNCHER(batch_launch_step)
LG_BATCH_LAUNCHER(batch_launch_diff)
LG_BATCH_LAUNCHER(batch_launch_solve)
#undef LG_BATCH_LAUNCHER
// the scalar kernel, with its moist, 3rd-order and diffusion-input variants chosen by tr, o3 and diff
template <class R>
void batch_launch_scalar(dim3 g, dim3 b, hipStream_t st, const MemberDev<R> *M, const int *prefix, int nm, const BatchArgs &a,
                         bool tr, bool o3, bool diff);

// the radiation phases (all on the PH_RAD grid)
enum RadPhase : int { RPH_SOLVER, RPH_GATHER, RPH_APPLY, RPH_SELECT };
template <class R>
void batch_launch_radiation(int rph, dim3 g, dim3 b, hipStream_t st, const MemberDev<R> *M, const int *prefix, int nm,
                            const BatchArgs &a);

}  // namespace lgame
