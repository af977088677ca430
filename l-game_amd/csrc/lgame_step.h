// lgame_step.h -- kernel declarations and the host-side step sequence.
#pragma once

#include "lgame_dev.h"

namespace lgame {

__global__ void k_temperature(const Ctx c, const StateDev s);
__global__ void k_theta_fill(const Ctx c, const StateDev s, int nfield);
__global__ void k_theta_grad(const Ctx c, int nfield);
__global__ void k_theta_div(const Ctx c, int nfield);
__global__ void k_theta_edge(const Ctx c, const double *wind_u, const double *wind_v, int nfield);
__global__ void k_pgrad(const Ctx c, const StateDev s, int lfirst);
__global__ void k_pot_vort(const Ctx c, const StateDev s);
__global__ void k_mass_flux(const Ctx c, const StateDev s);
__global__ void k_vort_flux(const Ctx c);
__global__ void k_ke(const Ctx c, const StateDev s);
__global__ void k_ke_grad(const Ctx c);
__global__ void k_dissipation(const Ctx c, const StateDev s);
__global__ void k_vec_tend(const Ctx c, const StateDev old, const StateDev nw, int rk);
__global__ void k_scalar_flux(const Ctx c, const StateDev ss, const StateDev nw);
__global__ void k_scalar_div(const Ctx c, const StateDev ss, const StateDev nw, int rk);
__global__ void k_solve_ver(const Ctx c, const StateDev old, const StateDev nw, int rk);
__global__ void k_solve_gen(const Ctx c, const StateDev old, const StateDev nw, int rk);

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

// Enqueue one HEVI predictor-corrector step (pchevi, mo_manage_pchevi.F90:27-118) on `stream`.
void enqueue_pchevi(const Ctx &c, const StateDev &old, const StateDev &nw, int total_step_counter, hipStream_t stream,
                    ProfHook *prof);

}  // namespace lgame
