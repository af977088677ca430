// lgame_dev.h -- device-side data layout and index helpers of the MI355X HEVI step.
//
// Layout in HBM: every 3-D field is stored level-major, then y, then x (x fastest), in one
// uniform padded box of (n_layers+1) levels x (ny+1) rows x pitch columns (pitch = nx+1 rounded
// up to 16 doubles = 128 B), so one wavefront reads 64 consecutive x points of a row and all
// staggered fields (u: nx+1 columns, v: ny+1 rows, w: n_layers+1 levels) share one indexing rule.
// Device functions take the reference's 1-based (ji, jk, jl) = (y, x, level) indices so the
// arithmetic reads like the reference loops it restates.
//
// Precision: the state is always FP64.  The grid metrics and the work arrays of the stencil
// kernels have the type R of the form: double in the plain form, float in the mixed form.  The
// column solvers read FP64 metrics (`g64`), so in the mixed form the grid exists in both types.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <type_traits>

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
// soil layers the surface kernel holds in registers
constexpr int MAX_SOIL = 16;

// XCDs of the MI355X; blocks of a launch go to them round-robin.
constexpr int N_XCD = 8;

// Element index inside one field.  Every field is smaller than 4 GiB (checked by lgame_init),
// so the byte offset fits 32 bits and an access compiles to one load with a scalar base
// register and a 32-bit vector offset.
using ix_t = unsigned;

// Pointer to one field; operator[] forms the 32-bit byte offset.
template <class T>
struct Fld {
  T *p = nullptr;
  Fld() = default;
  __host__ __device__ Fld(T *q) : p(q) {}
  template <class U>
  __host__ __device__ Fld(const Fld<U> &o) : p(o.p) {}
  __device__ __forceinline__ T &operator[](ix_t i) const {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(const_cast<typename std::remove_const<T>::type *>(p)) +
                                  i * (ix_t)sizeof(T));
  }
};

struct Geom {
  int ny, nx, nl, nlev, nflat, nc, ncond;
  int pitch;       // row pitch in elements
  long long lvl;   // elements per level = (ny+1)*pitch
  long long fsize; // elements per field = nlev*lvl
};

// Read-only grid metrics in the device layout.
template <class R>
struct GridT {
  Fld<const R> z_scalar, z_u, z_v, z_w, dx, dy, dz, area_x, area_y, area_z, volume;
  Fld<const R> slope_x, slope_y, ipw[6], area_dual_x, area_dual_y, area_dual_z;
  Fld<const R> z_area_dual_z, fvec_x, fvec_y, fvec_z, ebg_u, ebg_v, ebg_w, theta_v_bg;
  Fld<const R> exner_bg, gravity_m_v, gravity_potential;
  Fld<const R> klemp_pref;  // 1 + damping_coeff*dtime at w levels (mo_column_solvers.F90:190-196)
  Fld<const R> dx_dual, dy_dual, layer_thickness;  // read by the diffusion kernels only
  Fld<const R> land_fraction, roughness_length;    // 2-D (level 1), read by the surface kernels only
  // 2-D surface fields of the soil at levels 1..5: lake_fraction, sfc_rho_c, t_conduc_soil,
  // t_const_soil and the sea surface temperature (diag%sst, constant over a run)
  Fld<const R> soil;
  // soil grid (FP64 in both forms): z_soil_interface(1:nsoillays+1), z_soil_center(1:nsoillays),
  // z_t_const, consecutively
  const double *soil_z;
  const R *trsk_u;          // (3, ny) as in the reference
  const R *trsk_v;          // (2, ny+1)
};
using GridDev = GridT<double>;

struct StateDev {
  Fld<double> rho[MAX_CONST];
  Fld<double> rhotheta_v, theta_v_pert, exner_pert, wind_u, wind_v, wind_w;
  Fld<double> temperature_soil;  // soil layer jl at level jl (with the prognostic soil)
};

// Tendencies, carried diagnostics and scratch of one handle.
template <class R>
struct WorkT {
  Fld<R> tend_rho[MAX_CONST];
  Fld<R> tend_rhotheta_v, tend_u, tend_v, tend_w;
  Fld<R> pnl_u, pnl_v, pnl_w, pl_u, pl_v, pl_w, pold_u, pold_v;
  Fld<R> vsg_x, vsg_y, vsg_z, pvt_x, pvt_y, pvt_z;
  Fld<R> heating_diss, csh, temperature, radiation_tendency;
  Fld<R> eta_x, eta_y, eta_z, u_ph, v_ph, w_ph, v_squared;
  Fld<R> tgu[2], tgv[2], tfd[2], tu[2], tv[2];  // theta_v_adv_3rd_order: full, pert
  Fld<R> decel, pgc;  // per cell of the old state: C_D_P*rho_gas/rho_sum and rho_gas/rho_sum - 1
  Fld<double> col[12];                                  // column-solver per-level scratch (FP64)
  Fld<double> vfall_cw;  // sink velocity of cloud water per layer (moist), with the temperature
  // diffusion and turbulence (allocated only when a diffusion switch is on).  Carried between
  // steps as the reference's t_diag carries them: tke, visc (density times the horizontal
  // kinematic viscosity, diag%viscosity), mom_x/_y/_z (diag%mom_diff_tend_*), tdh
  // (diag%temp_diff_heating) and mdt (diag%mass_diff_tendency).  Per step: zeta (relative
  // vorticity at the dual points), tke_next, kv (molecular plus vertical eddy viscosity of a
  // cell), vdiv (visc times the horizontal divergence), vz (viscosity_dual times zeta) and the
  // diffusive fluxes dfu, dfv, dfw of the temperature (index 0) and the densities (1..nc).
  // rsum is sum(rho(1:ncond+1)) of the old state, stored by k_temperature.
  Fld<R> tke, tke_next, visc, kv, vdiv, zeta, vz, mom_x, mom_y, mom_z, tdh, mdt[MAX_CONST], rsum;
  Fld<R> dfu[MAX_CONST + 1], dfv[MAX_CONST + 1], dfw[MAX_CONST + 1];
  // surface turbulence: diag%roughness_length, roughness_velocity, monin_obukhov_length at
  // levels 1, 2, 3; surface heat: diag%scalar_flux_resistance, power_flux_density_sens_soil and
  // power_flux_density_sens_sea at levels 1, 2, 3 of sfc2
  Fld<R> sfc, sfc2;
};

struct Cfg {
  double dtime, toa, dxn, dyn, klemp_damp_max, klemp_begin_rel;
  int order, llinear, lcorio, lmoist, lklemp, lfreeslip, n_damping_levels;
  // diffusion (mo_diff_nml.F90); eff_hor_res = sqrt(dx*dy) (mo_run_nml.F90:93), mva =
  // grid%mean_velocity_area (mo_grid_generator.F90:1111)
  int lmom_h, lmom_v, ltemp_h, ltemp_v, lmass_h, lmass_v, scheme_tke, lpbl;
  double c_s, bg_shear, eff_hor_res, mva, h_prandtl, karman;
  // surface (mo_surface_nml.F90): prognostic soil with nsoil layers, sensible heat flux, and
  // whether the surface turbulence quantities are updated (lpbl, the sensible heat flux or the
  // surface phase transitions)
  int lsoil, lsens, lsfc_turb, nsoil;
  // the surface shortwave input and longwave output of the last radiation update (device column
  // order (ji-1)*nx + (jk-1)), read by the soil; null without radiation
  const double *rad_sw_in, *rad_lw_out;
  // set (to 1) by the corrector's column solvers when a new state value is not finite
  unsigned long long *nonfinite;
};

template <class R>
struct CtxT {
  Geom G;
  Cfg C;
  GridT<R> g;         // metrics read by the stencil kernels
  GridT<double> g64;  // metrics read by the column solvers and budget sums
  WorkT<R> w;
};
using Ctx = CtxT<double>;

#define LG_IX(G, ji, jk, jl) \
  ((ix_t)((jl)-1) * (ix_t)(G).lvl + (ix_t)((ji)-1) * (ix_t)(G).pitch + (ix_t)((jk)-1))

}  // namespace lgame
