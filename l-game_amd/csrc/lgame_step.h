// lgame_step.h -- kernel declarations and the host-side step sequence.
#pragma once

#include "lgame_dev.h"
#include "lgame_rad.h"

namespace lgame {

template <class R> __global__ void k_temperature(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_theta_grad(const CtxT<R> c, const StateDev s, int nfield);
template <class R> __global__ void k_theta_div(const CtxT<R> c, int nfield);
template <class R> __global__ void k_theta_edge(const CtxT<R> c, const StateDev s, int nfield);
// the potential vorticity, the mass fluxes and (with_ke) the kinetic energy
template <class R> __global__ void k_pot_vort(const CtxT<R> c, const StateDev s, int with_ke);
template <class R> __global__ void k_pot_vort2(const CtxT<R> c, const StateDev s, int with_ke);  // nx even: two cells per thread
template <class R> __global__ void k_vort_flux(const CtxT<R> c, int with_ke_grad);
template <class R> __global__ void k_dissipation(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_vec_tend(const CtxT<R> c, const StateDev old, const StateDev nw, int rk);
// the first substep's pressure gradient (manage_p_grad) with its wind update (the work of k_vec_tend)
template <class R> __global__ void k_pgrad_vt(const CtxT<R> c, const StateDev old, const StateDev nw, int lfirst);
// the second substep's vorticity flux term (and kinetic-energy gradient) with its wind update
template <class R> __global__ void k_vort_flux_vt(const CtxT<R> c, const StateDev old, const StateDev nw, int with_ke_grad);
template <class R> __global__ void k_solve_ver(const CtxT<R> c, const StateDev old, const StateDev nw, int rk);
template <class R> __global__ void k_solve_gen(const CtxT<R> c, const StateDev old, const StateDev nw, int rk);
template <class R> __global__ void k_diff_zeta(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_diff_coef(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_diff_coef2(const CtxT<R> c, const StateDev s);  // nx even: two cells per thread
template <class R> __global__ void k_diff_dual(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_diff_mom(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_diff_flux(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_diff_flux2(const CtxT<R> c, const StateDev s);  // nx even: two cells per thread
template <class R> __global__ void k_diff_div(const CtxT<R> c);
template <class R> __global__ void k_diff_div2(const CtxT<R> c);  // nx even: two cells per thread
template <class R> __global__ void k_sfc_turb(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_pbl_wind(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_surface(const CtxT<R> c, const StateDev old, const StateDev nw, int rk);

// Launch geometry of the horizontal stencil kernels: 64 x-points by BY rows per block (the
// default; the tile table can choose 1, 2 or 4 rows per kernel).
constexpr int BX = 64;
constexpr int BY = 4;

// stencil kernels of the step, the index of their tile-table entry.  The mass fluxes and the
// kinetic energy are formed in the potential-vorticity kernel, the kinetic-energy gradient in the
// vorticity-flux kernel, the scalar fluxes in the scalar column kernel and the full theta_v of the
// third-order theta where the gradient and edge kernels load it: KI_THETA_FILL, KI_MASS_FLUX, KI_KE,
// KI_KE_GRAD and KI_SCALAR_FLUX keep their place in the table (and the C-ABI's order) unused.
enum KernelId {
  KI_TEMPERATURE, KI_THETA_FILL, KI_THETA_GRAD, KI_THETA_DIV, KI_THETA_EDGE, KI_PGRAD, KI_POT_VORT, KI_MASS_FLUX,
  KI_VORT_FLUX, KI_KE, KI_KE_GRAD, KI_DISSIPATION, KI_VEC_TEND, KI_SCALAR_FLUX, KI_SCALAR_DIV, KI_DIFF_ZETA,
  KI_DIFF_COEF, KI_DIFF_DUAL, KI_DIFF_MOM, KI_DIFF_FLUX, KI_DIFF_DIV, KI_COUNT
};

// rows per block of every stencil kernel (block = 64 x by threads), and the number of vertical
// segments of the column-marching scalar kernel
struct Tiles {
  int by[KI_COUNT];
  int seg = 1;
  int ver_passes = 1;  // the vertical solver in one column pass (k_solve_ver) or three (launch_solve_ver)
  Tiles() {
    for (int &b : by) b = BY;
  }
};

// The column-marching scalar_tend_expl (k_scalar_col): 64-lane rows of which 62 lanes are cells
// and two the west and east halo; `rows` rows per block, the levels cut into `nseg` segments.
constexpr int SCOL_CELLS = 62;
// segments that give a domain about four wavefronts per SIMD of the MI355X (1024 SIMDs), with at
// least four levels per segment
inline int default_segments(const Geom &G) {
  const long long waves = (long long)((G.nx + SCOL_CELLS - 1) / SCOL_CELLS) * G.ny;
  int sg = 1;
  while (waves * sg < 4096 && sg < 8 && G.nl / (2 * sg) >= 4) sg *= 2;
  return sg;
}
template <class R>
void launch_scalar_col(const CtxT<R> &c, const StateDev &ss, const StateDev &nw, int rk, int rows, int nseg, hipStream_t st);
// Lanes per block of the column solvers.
constexpr int BCOL = 64;
// The vertical solver runs in three passes (rows, Thomas recursion per column, new state) on
// domains with fewer column wavefronts than this, the 1024 SIMDs of the MI355X.
constexpr int VER_SPLIT_WAVES = 1024;
template <class R>
void launch_solve_ver(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int rk, int passes, hipStream_t st);
// The column solvers give each wavefront 64 consecutive columns of one row (the row's last
// wavefront partly idle), so that a wavefront's loads of a level cover whole 128-byte lines of
// the row rather than the end of one row and the start of the next.
__host__ __device__ inline int col_waves_per_row(int nx) { return (nx + BCOL - 1) / BCOL; }
__host__ __device__ inline int col_waves(const Geom &G) { return col_waves_per_row(G.nx) * G.ny; }
// The tracer solver: one wavefront per tracer of a 64-column group, two groups per block. Twelve
// wavefronts per block land three on each SIMD, which is what the kernel's registers allow; a
// block of six would land 2,2,1,1 and leave room for only one block per CU.
constexpr int NTR = 6;         // tracers per moist column
constexpr int GEN_GROUPS = 2;  // 64-column groups per block

// Optional timing hook around kernel groups (0 = vertical solver, 1 = tracer solver, 2 = step,
// 3 = radiation; PROF_KERNEL0 + KernelId for every stencil kernel when `kernels` is set).
constexpr int PROF_KERNEL0 = 8;
constexpr int PROF_GROUPS = PROF_KERNEL0 + KI_COUNT;
struct ProfHook {
  bool kernels = false;
  virtual void begin(int group) = 0;
  virtual void end(int group) = 0;
  virtual ~ProfHook() = default;
};

// Enqueue one HEVI predictor-corrector step (pchevi, mo_manage_pchevi.F90:27-118) on `stream`;
// `rad` is non-null when the step updates the radiation (lrad_update).  Returns 0, or the
// LGAME_E* code of a radiation update that could not be enqueued.
template <class R>
int enqueue_pchevi(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int total_step_counter, hipStream_t stream,
                    ProfHook *prof, const RadCall *rad, const Tiles &tiles);

}  // namespace lgame
