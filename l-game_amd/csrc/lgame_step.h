// lgame_step.h -- kernel declarations and the host-side step sequence.
#pragma once

#include "lgame_dev.h"
#include "lgame_rad.h"

namespace lgame {

template <class R> __global__ void k_temperature(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_theta_fill(const CtxT<R> c, const StateDev s, int nfield);
template <class R> __global__ void k_theta_grad(const CtxT<R> c, int nfield);
template <class R> __global__ void k_theta_div(const CtxT<R> c, int nfield);
template <class R> __global__ void k_theta_edge(const CtxT<R> c, Fld<const double> wind_u, Fld<const double> wind_v, int nfield);
template <class R> __global__ void k_pgrad(const CtxT<R> c, const StateDev s, int lfirst);
template <class R> __global__ void k_pot_vort(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_mass_flux(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_vort_flux(const CtxT<R> c);
template <class R> __global__ void k_ke(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_ke_grad(const CtxT<R> c);
template <class R> __global__ void k_dissipation(const CtxT<R> c, const StateDev s);
template <class R> __global__ void k_vec_tend(const CtxT<R> c, const StateDev old, const StateDev nw, int rk);
template <class R> __global__ void k_scalar_flux(const CtxT<R> c, const StateDev ss, const StateDev nw);
template <class R> __global__ void k_scalar_div(const CtxT<R> c, const StateDev ss, const StateDev nw, int rk);
template <class R> __global__ void k_solve_ver(const CtxT<R> c, const StateDev old, const StateDev nw, int rk);
template <class R> __global__ void k_solve_gen(const CtxT<R> c, const StateDev old, const StateDev nw, int rk);

// Launch geometry of the horizontal stencil kernels: 64 x-points by 4 rows per block.
constexpr int BX = 64;
constexpr int BY = 4;
// Lanes per block of the column solvers.
constexpr int BCOL = 64;

// Optional timing hook around kernel groups (0 = vertical solver, 1 = tracer solver).
struct ProfHook {
  virtual void begin(int group) = 0;
  virtual void end(int group) = 0;
  virtual ~ProfHook() = default;
};

// Enqueue one HEVI predictor-corrector step (pchevi, mo_manage_pchevi.F90:27-118) on `stream`;
// `rad` is non-null when the step updates the radiation (lrad_update).
template <class R>
void enqueue_pchevi(const CtxT<R> &c, const StateDev &old, const StateDev &nw, int total_step_counter, hipStream_t stream,
                    ProfHook *prof, const RadCall *rad);

}  // namespace lgame
