// lgame_dev.h -- device-side data layout and index helpers of the MI355X HEVI step.
//
// Layout in HBM: every 3-D field is stored level-major, then y, then x (x fastest), in one
// uniform padded box of (n_layers+1) levels x (ny+1) rows x pitch columns (pitch = nx+1 rounded
// up to 16 doubles = 128 B), so one wavefront reads 64 consecutive x points of a row and all
// staggered fields (u: nx+1 columns, v: ny+1 rows, w: n_layers+1 levels) share one indexing rule.
// Device functions take the reference's 1-based (ji, jk, jl) = (y, x, level) indices so the
// arithmetic reads like the reference loops it restates.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace lgame {

// physical constants, src/mo_constants.F90:16-36
constexpr double R_E = 6371000.789927;
constexpr double N_A = 6.02214076e23;
constexpr double T_0 = 273.15;
constexpr double RHO_H2O = 1024.0;
constexpr double R_D = 287.057811;
constexpr double C_D_V = 717.942189;
constexpr double C_V_V = 1396.475121;
constexpr double C_D_P = 1005.0;
constexpr double K_B = 1.380649e-23;
constexpr double M_D = N_A * 0.004810e-23;
constexpr double M_V = N_A * 0.002991e-23;
constexpr double IMPL_THERMO_WEIGHT = C_D_V / C_D_P;
constexpr double MPI = 3.14159265358979323846;

constexpr int MAX_CONST = 7;

struct Geom {
  int ny, nx, nl, nlev, nflat, nc, ncond;
  int pitch;       // row pitch in elements
  long long lvl;   // elements per level = (ny+1)*pitch
  long long fsize; // elements per field = nlev*lvl
};

// Read-only grid metrics in the device layout.
struct GridDev {
  const double *z_scalar, *z_u, *z_v, *z_w, *dx, *dy, *dz, *area_x, *area_y, *area_z, *volume;
  const double *slope_x, *slope_y, *ipw[6], *area_dual_x, *area_dual_y, *area_dual_z;
  const double *z_area_dual_z, *fvec_x, *fvec_y, *fvec_z, *ebg_u, *ebg_v, *ebg_w, *theta_v_bg;
  const double *exner_bg, *gravity_m_v, *gravity_potential;
  const double *klemp_pref;  // 1 + damping_coeff*dtime at w levels (mo_column_solvers.F90:190-196)
  const double *trsk_u;      // (3, ny) as in the reference
  const double *trsk_v;      // (2, ny+1)
};

struct StateDev {
  double *rho[MAX_CONST];
  double *rhotheta_v, *theta_v_pert, *exner_pert, *wind_u, *wind_v, *wind_w;
};

// Tendencies, carried diagnostics and scratch of one handle.
struct WorkDev {
  double *tend_rho[MAX_CONST];
  double *tend_rhotheta_v, *tend_u, *tend_v, *tend_w;
  double *pnl_u, *pnl_v, *pnl_w, *pl_u, *pl_v, *pl_w, *pold_u, *pold_v;
  double *vsg_x, *vsg_y, *vsg_z, *pvt_x, *pvt_y, *pvt_z;
  double *heating_diss, *csh, *temperature, *radiation_tendency;
  double *eta_x, *eta_y, *eta_z, *u_ph, *v_ph, *w_ph, *v_squared;
  double *tth[2], *tgu[2], *tgv[2], *tfd[2], *tu[2], *tv[2];  // theta_v_adv_3rd_order: full, pert
  double *fu, *fv, *thu, *thv;                       // gas mass flux and theta_v flux
  double *col[14];  // column-solver per-level scratch
};

struct Cfg {
  double dtime, toa, dxn, dyn, klemp_damp_max, klemp_begin_rel;
  int order, llinear, lcorio, lmoist, lklemp, lfreeslip, n_damping_levels;
};

struct Ctx {
  Geom G;
  Cfg C;
  GridDev g;
  WorkDev w;
};

#define LG_IX(G, ji, jk, jl) \
  ((size_t)((jl)-1) * (size_t)(G).lvl + (size_t)((ji)-1) * (size_t)(G).pitch + (size_t)((jk)-1))

}  // namespace lgame
