/*
 * lgame_gpu.h -- C-ABI of the MI355X HEVI predictor-corrector step.
 *
 * This is the drop-in boundary that replaces the reference time-stepping entry
 *   subroutine pchevi(state_old,state_new,tend,bc,grid,diag,total_step_counter,lrad_update,t_0)
 *   (src/time_stepping/mo_manage_pchevi.F90:27-118), called from the serial time loop at
 *   src/control.F90:520-524 with the two state slots swapped on odd steps.
 *
 * Everything crossing the boundary is plain C: ints, doubles and pointers to host arrays that
 * keep the reference's own layout (Fortran column-major, 1-based in Fortran, y fastest):
 *   scalar fields  (ny,   nx,   n_layers)          e.g. rhotheta_v, exner_pert
 *   u-points       (ny,   nx+1, n_layers)          wind_u, dx, area_x, slope_x, ...
 *   v-points       (ny+1, nx,   n_layers)          wind_v, dy, area_y, slope_y, ...
 *   w-points       (ny,   nx,   n_layers+1)        wind_w, dz, z_w, area_z, ...
 *   rho            (ny,   nx,   n_layers, n_constituents)
 *   inner_product_weights (6, ny, nx, n_layers)    (6 is the fastest index)
 * A Fortran caller passes c_loc() of its allocatable arrays (allocatables are contiguous).
 *
 * The module globals the reference step reads (mo_run_nml, mo_constituents_nml, mo_diff_nml,
 * mo_surface_nml, mo_bc_nml, mo_rad_nml) become the lgame_dims / lgame_cfg structs below;
 * Fortran LOGICALs cross as int (0 = .false., nonzero = .true.).
 *
 * Error convention: every entry point returns 0 on success and a nonzero LGAME_E* code
 * otherwise; nothing is thrown across the ABI.  lgame_last_error() gives the message.  The
 * reference aborts in place with "call exit(1)" (mo_run_nml.F90:104-141, mo_write_out.F90:63-67);
 * a Fortran shim keeps that by printing lgame_last_error() and calling exit(1) on nonzero.
 *
 * Threading: one host thread per handle; a handle is not re-entrant.  Each handle owns one HIP
 * stream; all device work of a handle is ordered on it.
 */
#ifndef LGAME_GPU_H
#define LGAME_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define LGAME_ABI_VERSION 4

/* return codes */
#define LGAME_OK          0
#define LGAME_EINVAL      1   /* bad argument or unsupported configuration   */
#define LGAME_EHIP        2   /* a HIP runtime call failed                    */
#define LGAME_ENOMEM      3   /* device allocation failed                     */
#define LGAME_ENONFINITE  4   /* NaN/Inf found in a prognostic field          */
#define LGAME_ESTATE      5   /* call out of order (e.g. step before upload)  */

/* precision modes */
#define LGAME_PREC_FP64   0   /* plain form: FP64 everywhere                                   */
#define LGAME_PREC_MIXED  1   /* FP32 stencils/explicit tendencies, FP64 solves/updates/sums   */

/*
 * Radiation modes (lgame_cfg.rad_mode; the step's lrad_update input only has an effect when not
 * OFF and lrad is set).  The reference computes radiative heating with RTE+RRTMGP
 * (src/radiation/mo_rterrtmgp_coupler.F90:56-493), a third-party library that is not part of the
 * reference tree.  This library keeps the coupler's own steps (the pressure and temperature
 * interfaces, the water paths, coszenith, calc_power_density and the surface fluxes) and puts a
 * gray two-stream solver in the place of the RTE+RRTMGP calls, with the constants below:
 *   longwave:  per-layer optical depth  dm*(K_LW_DRY*p/P_0 + K_LW_VAPOUR*q_v) + K_LW_LIQUID*LWP
 *              + K_LW_ICE*IWP, transmissivity exp(-DIFFUSIVITY*tau), layer source sigma*T^4,
 *              surface emission EMISSIVITY*sigma*T_sfc^4 plus the reflected rest;
 *   shortwave: direct beam SOLAR_CONSTANT*mu_0 attenuated by exp(-tau_sw/mu_0), tau_sw =
 *              dm*(K_SW_DRY + K_SW_VAPOUR*q_v) + K_SW_LIQUID*LWP + K_SW_ICE*IWP, reflected at the
 *              surface with the albedo and attenuated upward by exp(-DIFFUSIVITY*tau_sw);
 * dm = layer air mass per area from the interface pressures [kg m^-2], q_v = rho_vapour/rho_gas,
 * LWP/IWP the reference's water paths [g m^-2].
 *   LGAME_RAD_TEACHER:   that solver on every column (FP64); it is also the trainer's teacher.
 *   LGAME_RAD_SURROGATE: the per-column MLP attached with lgame_rad_attach_mlp; each call also
 *                        re-runs the solver on an audit sample of the columns and gives every
 *                        audited column whose heating-rate RMSE exceeds the tolerance the
 *                        solver's result (lgame_rad_stats counts both).  A column whose
 *                        predicted outputs leave the network's training envelope (any output
 *                        more than LGAME_RAD_ENVELOPE standard deviations of its training
 *                        targets from their mean, or not finite) is always given the solver's
 *                        result.
 */
#define LGAME_RAD_ENVELOPE   6.0
#define LGAME_RAD_OFF        0
#define LGAME_RAD_TEACHER    1
#define LGAME_RAD_SURROGATE  2

#define LGAME_RAD_SOLAR_CONSTANT  1361.0      /* W m^-2                                         */
#define LGAME_RAD_SCALE_HEIGHT    8.e3        /* m, mo_rterrtmgp_coupler.F90:126                */
#define LGAME_RAD_T_MIN           160.0       /* K, temperature clipping of the coupler          */
#define LGAME_RAD_T_MAX           355.0       /* K                                               */
#define LGAME_RAD_EMISSIVITY      0.98        /* mo_rterrtmgp_coupler.F90:150                    */
#define LGAME_RAD_DIFFUSIVITY     1.66
#define LGAME_RAD_P_0             100000.0    /* Pa                                              */
#define LGAME_RAD_GRAVITY         9.80616     /* m s^-2                                          */
#define LGAME_RAD_SIGMA           5.670374419e-8
#define LGAME_RAD_K_LW_DRY        2.0e-4      /* m^2 kg^-1 (times p/P_0)                         */
#define LGAME_RAD_K_LW_VAPOUR     0.08        /* m^2 kg^-1                                       */
#define LGAME_RAD_K_LW_LIQUID     0.13        /* m^2 g^-1                                        */
#define LGAME_RAD_K_LW_ICE        0.07        /* m^2 g^-1                                        */
#define LGAME_RAD_K_SW_DRY        1.0e-5      /* m^2 kg^-1                                       */
#define LGAME_RAD_K_SW_VAPOUR     0.02        /* m^2 kg^-1                                       */
#define LGAME_RAD_K_SW_LIQUID     0.02        /* m^2 g^-1                                        */
#define LGAME_RAD_K_SW_ICE        0.01        /* m^2 g^-1                                        */

/*
 * Surrogate column interface (LGAME_RAD_LEVELS = 60 levels): inputs are 5 profiles of 60 levels
 * (temperature [K], pressure [Pa], vapour, liquid and ice mixing ratios [kg/kg]; liquid = rain +
 * cloud water, ice = snow + cloud ice + graupel) followed by 4 scalars (surface temperature [K],
 * mu_0, surface albedo, surface pressure [Pa]); outputs are the 60 radiative heating rates [K/day]
 * followed by sfc_sw_in and sfc_lw_out [W m^-2].  A domain of 60 layers maps them one to one;
 * another layer count is resampled linearly in the layer index (lgame_rad_attach_mlp).
 */
#define LGAME_RAD_LEVELS     60
#define LGAME_RAD_N_IN       (5 * LGAME_RAD_LEVELS + 4)
#define LGAME_RAD_N_OUT      (LGAME_RAD_LEVELS + 2)

/*
 * Dimensions.  Mirrors mo_run_nml.F90:13-19 (ny, nx, n_layers, n_oro_layers; n_levels and
 * n_flat_layers are derived as in mo_run_nml.F90:89-90), mo_constituents_nml.F90:12-15 and
 * mo_surface_nml.F90 (nsoillays).
 */
typedef struct lgame_dims {
  int ny;
  int nx;
  int n_layers;
  int n_oro_layers;
  int n_constituents;            /* 1 dry, 7 moist                                   */
  int n_condensed_constituents;  /* 0 dry, 5 moist                                   */
  int nsoillays;
} lgame_dims;

/*
 * Step configuration.  Names and meaning follow the reference namelist variables:
 * mo_run_nml.F90:13-41, mo_diff_nml.F90, mo_surface_nml.F90, mo_bc_nml.F90, mo_rad_nml.F90.
 * dtime is passed explicitly (the reference derives it from dx, dy at mo_run_nml.F90:93-94).
 *
 * Diffusion (mo_diff_nml.F90:8-53; the reference default turns all six switches on, with
 * diff_coeff_scheme_h "smag", diff_coeff_scheme_v "tke", c_s 0.2, bg_shear 1.5e-5): every
 * combination is accepted in which horizontal temperature or mass diffusion comes with horizontal
 * momentum diffusion.  Without it, scalar_diff_coeffs (mo_eff_diff_coeffs.F90:540-590) evaluates
 * the Smagorinsky coefficient from whatever vorticity and divergence the diagnostics last held and
 * updates the TKE a second time; lgame_init rejects those combinations with LGAME_EINVAL.
 * diff_coeff_scheme_v has one valid value in the reference ("tke", mo_diff_nml.F90:95-98), so it
 * has no field here.
 *
 * lpbl (mo_surface_nml.F90) turns on the surface turbulence quantities (update_sfc_turb_quantities,
 * mo_pbl.F90:19-109) and the PBL friction of the lowest-layer wind (pbl_wind_tendency,
 * :111-250), with h_prandtl and karman (mo_diff_nml.F90:39-44, defaults 100 m and 0.4); as in
 * the reference it needs lmom_diff_h.  lprog_soil_temp adds the soil temperature rows to the
 * vertical solve (mo_column_solvers.F90:229-336) with nsoillays in 2..16 layers;
 * lsfc_sensible_heat_flux adds the sensible heat flux over soil and sea to the lowest layer
 * (:81-120) and needs lprog_soil_temp (the scalar flux resistance it divides by exists only with
 * the soil, mo_pbl.F90:96-104).  lsfc_phase_trans only turns on the surface turbulence
 * quantities: the reference's surface phase transitions sit in calc_h2otracers_source_rates
 * (mo_phase_trans.F90:22-330), which nothing in the reference calls.  With
 * the prognostic soil and radiation, the radiation's surface temperature is the reference's mix
 * of the sea surface temperature and the top soil layer (mo_manage_radiation_calls.F90:36-42,
 * the temp_sfc of lgame_rad_view is not read) and the soil takes the surface short- and
 * longwave fluxes of the last radiation update.
 */
typedef struct lgame_cfg {
  double dtime;
  double toa;
  double dx;                 /* namelist dx: used by theta_v_adv_3rd_order (mo_multiplications.F90:142) */
  double dy;                 /* namelist dy                                                            */
  double klemp_damp_max;
  double klemp_begin_rel;
  double dtime_rad;
  double c_s;                /* Smagorinsky parameter                                                  */
  double bg_shear;           /* minimum background shear [1/s]                                         */
  double h_prandtl;          /* height of the Prandtl layer [m] (PBL)                                  */
  double karman;             /* von Karman's constant (PBL)                                            */
  int theta_adv_order;       /* 2 or 3                                                                 */
  int luse_bg_state;
  int llinear;
  int lcorio;
  int lmoist;
  int lklemp;
  int lmom_diff_h, lmom_diff_v;
  int ltemp_diff_h, ltemp_diff_v;
  int lmass_diff_h, lmass_diff_v;
  int lprog_soil_temp;
  int lsfc_sensible_heat_flux;
  int lsfc_phase_trans;
  int lpbl;
  int lperiodic;
  int lfreeslip;
  int lrad;
  int precision;             /* LGAME_PREC_*                                                           */
  int rad_mode;              /* LGAME_RAD_*                                                            */
  int diff_coeff_scheme_h;   /* 0 = "smag", 1 = "tke"                                                  */
} lgame_cfg;

/*
 * Read-only grid metrics consumed by the step (the t_grid components of
 * src/mo_definitions.F90:29-91 that the hot path reads), in the reference layout.
 */
typedef struct lgame_grid_view {
  const double *z_scalar;              /* (ny,nx,nl)     */
  const double *z_u;                   /* (ny,nx+1,nl)   */
  const double *z_v;                   /* (ny+1,nx,nl)   */
  const double *z_w;                   /* (ny,nx,nl+1)   */
  const double *dx;                    /* (ny,nx+1,nl)   */
  const double *dy;                    /* (ny+1,nx,nl)   */
  const double *dz;                    /* (ny,nx,nl+1)   */
  const double *area_x;                /* (ny,nx+1,nl)   */
  const double *area_y;                /* (ny+1,nx,nl)   */
  const double *area_z;                /* (ny,nx,nl+1)   */
  const double *volume;                /* (ny,nx,nl)     */
  const double *slope_x;               /* (ny,nx+1,nl)   */
  const double *slope_y;               /* (ny+1,nx,nl)   */
  const double *inner_product_weights; /* (6,ny,nx,nl)   */
  const double *area_dual_x;           /* (ny+1,nx,nl+1) */
  const double *area_dual_y;           /* (ny,nx+1,nl+1) */
  const double *area_dual_z;           /* (ny+1,nx+1,nl) */
  const double *z_area_dual_z;         /* (ny+1,nx+1,nl) */
  const double *fvec_x;                /* (ny+1,nx)      */
  const double *fvec_y;                /* (ny,nx+1)      */
  const double *fvec_z;                /* (ny+1,nx+1)    */
  const double *trsk_weights_u;        /* (3,ny)         */
  const double *trsk_weights_v;        /* (2,ny+1)       */
  const double *exner_bg_grad_u;       /* (ny,nx+1,nl)   */
  const double *exner_bg_grad_v;       /* (ny+1,nx,nl)   */
  const double *exner_bg_grad_w;       /* (ny,nx,nl+1)   */
  const double *theta_v_bg;            /* (ny,nx,nl)     */
  const double *exner_bg;              /* (ny,nx,nl)     */
  const double *gravity_m_v;           /* (ny,nx,nl+1)   */
  const double *gravity_potential;     /* (ny,nx,nl)     */
  const double *dx_dual;               /* (ny+1,nx,nl)   read only with diffusion */
  const double *dy_dual;               /* (ny,nx+1,nl)   read only with diffusion */
  const double *layer_thickness;       /* (ny,nx,nl)     read only with diffusion */
  const double *land_fraction;         /* (ny,nx)        read only with the surface switches */
  const double *roughness_length;      /* (ny,nx)        read only with the surface switches */
  const double *lake_fraction;         /* (ny,nx)        read only with lprog_soil_temp     */
  const double *sfc_rho_c;             /* (ny,nx)        volumetric heat capacity, same     */
  const double *t_conduc_soil;         /* (ny,nx)        temperature conductivity, same     */
  const double *t_const_soil;          /* (ny,nx)        temperature at z_t_const, same     */
  const double *sst;                   /* (ny,nx)        diag%sst, constant over a run, same */
  const double *z_soil_interface;      /* (nsoillays+1)  same                               */
  const double *z_soil_center;         /* (nsoillays)    same                               */
  double mean_velocity_area;           /* grid%mean_velocity_area, mo_grid_generator.F90:1111 */
  double z_t_const;                    /* grid%z_t_const, mo_grid_generator.F90:1363          */
} lgame_grid_view;

/* Prognostic state (t_state, src/mo_definitions.F90:94-105), reference layout. */
typedef struct lgame_state_view {
  double *rho;               /* (ny,nx,nl,n_constituents) */
  double *rhotheta_v;        /* (ny,nx,nl)                */
  double *theta_v_pert;      /* (ny,nx,nl)                */
  double *exner_pert;        /* (ny,nx,nl)                */
  double *wind_u;            /* (ny,nx+1,nl)              */
  double *wind_v;            /* (ny+1,nx,nl)              */
  double *wind_w;            /* (ny,nx,nl+1)              */
  double *temperature_soil;  /* (ny,nx,nsoillays); read and written with lprog_soil_temp, else may be NULL */
} lgame_state_view;

typedef struct lgame_handle lgame_handle;

/*
 * Per-column surface and location data of the radiation call (update_rad_fluxes,
 * src/radiation/mo_manage_radiation_calls.F90:17-59): geographic latitude and longitude [rad],
 * surface albedo and the surface temperature the coupler receives (temp_sfc_full), each (ny,nx).
 */
typedef struct lgame_rad_view {
  const double *lat_geo;
  const double *lon_geo;
  const double *sfc_albedo;
  const double *temp_sfc;
} lgame_rad_view;

/* Budget-sum slots written by lgame_integrals (src/io/mo_write_out.F90:226-304). */
#define LGAME_INT_RHOTHETA   0   /* sum(rhotheta_v*volume)                                  */
#define LGAME_INT_KINETIC    1   /* 0.5*sum(rho_g*(ipw-weighted |v|^2)*volume)              */
#define LGAME_INT_POTENTIAL  2   /* sum(rho_g*gravity_potential*volume)                     */
#define LGAME_INT_INTERNAL   3   /* sum(c_d_v*rho_g*temperature*volume)                     */
#define LGAME_INT_MASS0      4   /* masses: slot LGAME_INT_MASS0 + jc-1 for constituent jc  */

/*
 * Checkpoint of the fields the step carries from one call to the next besides the two state
 * slots: the pressure-gradient terms of the last first substep (their negated sum is the next
 * step's old gradient, mo_p_grad.F90:38-43), the lagged momentum advection and the kinetic energy
 * (computed only at the corrector, mo_vector_tend_expl.F90:58; the surface turbulence of the next
 * step reads the latter, mo_pbl.F90:19-109), the sediment heating, the dissipative heating (the next
 * step's TKE production, mo_tke.F90:19-67), the radiation tendency and the
 * surface fluxes of the last radiation update with the count of updates (which picks the audit
 * sample), and with diffusion or the surface physics the TKE, the viscosity, the momentum-
 * diffusion tendencies, the diffusive heating and mass tendencies and the surface turbulence
 * quantities.  Saved together with the two states (lgame_get_state) and loaded into a handle of
 * the same configuration (lgame_set_state, lgame_checkpoint_load), a run continues bit for bit at
 * the next step counter.  The blob is opaque (device layout and precision of the handle) and
 * checked on load (LGAME_EINVAL for another configuration or a short buffer).
 * No counterpart in the reference: its restart (restart, mo_set_initial_state.F90:259, called
 * at control.F90:470-471) reads the state only and starts the time loop at counter 0 again
 * (control.F90:506), so it recomputes the old pressure gradient and the advection from the
 * restart state; callers that need a restart to reproduce an uninterrupted run use this.
 */
int lgame_checkpoint_size(lgame_handle *h, long long *bytes);
int lgame_checkpoint_save(lgame_handle *h, void *buf, long long bytes);
int lgame_checkpoint_load(lgame_handle *h, const void *buf, long long bytes);

/* Diagnostic fields readable with lgame_get_diag (reference t_diag components). */
#define LGAME_DIAG_TEMPERATURE     0   /* diag%temperature (ny,nx,nl), of the last step's state_old */
#define LGAME_DIAG_TKE             1   /* diag%tke (ny,nx,nl)                                       */
#define LGAME_DIAG_VISCOSITY       2   /* diag%viscosity (ny,nx,nl), density times the coefficient  */
#define LGAME_DIAG_MOM_DIFF_TEND_X 3   /* diag%mom_diff_tend_x (ny,nx+1,nl)                         */
#define LGAME_DIAG_MOM_DIFF_TEND_Z 4   /* diag%mom_diff_tend_z (ny,nx,nl+1)                         */
#define LGAME_DIAG_TEMP_DIFF_HEATING 5 /* diag%temp_diff_heating (ny,nx,nl)                         */
#define LGAME_DIAG_HEATING_DISS    6   /* diag%heating_diss (ny,nx,nl)                              */
#define LGAME_DIAG_ROUGHNESS_LENGTH 7  /* diag%roughness_length (ny,nx), with the surface switches   */
#define LGAME_DIAG_MONIN_OBUKHOV   8   /* diag%monin_obukhov_length (ny,nx), same                   */
#define LGAME_DIAG_SCALAR_FLUX_RESISTANCE 9 /* diag%scalar_flux_resistance (ny,nx), with the sensible heat flux */
#define LGAME_DIAG_PFD_SENS_SOIL   10  /* diag%power_flux_density_sens_soil (ny,nx), same           */
#define LGAME_DIAG_PFD_SENS_SEA    11  /* diag%power_flux_density_sens_sea (ny,nx), same            */

/*
 * Grid generation on the device (grid_setup and bg_setup, src/io/mo_grid_generator.F90:29-1457,
 * for the analytic orographies of :326-361 and periodic boundaries).  Parameters mirror the
 * &run and &surface namelist entries the generator reads (mo_run_nml.F90, mo_surface_nml.F90);
 * orography_id 5 is a set-up extension: a sum of Gaussian hills (x0, y0, height, half width in
 * metres, in `hills`, 4 per hill) on the periodic domain.  orography_id 1 is the reference's
 * real orography (:336-616) with the orography interpolated to the grid given in `oro` (the
 * reference reads it from ETOPO files, which it does not ship): the device smooths it with
 * smooth_hor_scalar2 (:1512-1557, inverse-distance weights over all other points, O((nx ny)^2))
 * and keeps a tenth of the detail with SLEVE (:605-616); the land fraction stays 0.
 */
typedef struct lgame_grid_params {
  double dx, dy, toa, stretching_parameter;
  double lat_center, lon_center;  /* rotation of the grid's pole [rad]                           */
  int orography_id;               /* 0 flat, 1 given (smoothed), 2 Schaer, 3, 4, 5 Gaussian hills */
  int lsleve, lplane, lfreeslip;
  int n_hills;
  const double *hills;            /* (4, n_hills), read with orography_id 5                     */
  const double *oro;              /* (ny,nx) orography [m], read with orography_id 1            */
} lgame_grid_params;

/*
 * Output of lgame_grid_setup: the arrays of lgame_grid_view in the same order and reference
 * layout, written by the library, then the geographic coordinates of the scalar points (used
 * by the radiation) and the two scalars.  The surface is set as the reference's generator sets
 * it without land (land and lake fraction 0, roughness length 0.08 m, water heat capacity and
 * conductivity, t_const_soil from the latitude) with a sea surface temperature of 280 K (the
 * initial soil temperature, mo_set_initial_state.F90:236-247).
 */
typedef struct lgame_grid_out {
  double *z_scalar, *z_u, *z_v, *z_w, *dx, *dy, *dz, *area_x, *area_y, *area_z, *volume, *slope_x, *slope_y;
  double *inner_product_weights, *area_dual_x, *area_dual_y, *area_dual_z, *z_area_dual_z;
  double *fvec_x, *fvec_y, *fvec_z, *trsk_weights_u, *trsk_weights_v;
  double *exner_bg_grad_u, *exner_bg_grad_v, *exner_bg_grad_w, *theta_v_bg, *exner_bg, *gravity_m_v, *gravity_potential;
  double *dx_dual, *dy_dual, *layer_thickness, *land_fraction, *roughness_length;
  double *lake_fraction, *sfc_rho_c, *t_conduc_soil, *t_const_soil, *sst, *z_soil_interface, *z_soil_center;
  double *lat_geo_scalar, *lon_geo_scalar;  /* (ny,nx) */
  double mean_velocity_area;
  double z_t_const;
} lgame_grid_out;

/*
 * Fill *out with the grid of dims and params, computed on device 0 and copied to the caller's
 * arrays (every pointer of *out must be valid).  Replaces grid_setup and bg_setup
 * (mo_grid_generator.F90:29-1457) for the analytic orographies.  Returns LGAME_EINVAL for
 * unsupported parameters or, as the reference stops, for levels that are not monotonic or
 * orography that reaches the lowest flat level (:936-951).
 */
int lgame_grid_setup(const lgame_dims *dims, const lgame_grid_params *params, lgame_grid_out *out);

/* Version of this ABI (LGAME_ABI_VERSION). */
int lgame_abi_version(void);

/*
 * Create a handle: validate dims/cfg, allocate device memory, upload the grid once and
 * precompute grid-only factors.  Replaces the set-up half of control.F90 that feeds pchevi.
 * Device 0 of the calling process's visible devices is used.
 */
int lgame_init(const lgame_dims *dims, const lgame_cfg *cfg, const lgame_grid_view *grid,
               lgame_handle **out);

/*
 * Host -> device copy of a full state into resident slot 0 or 1 (lgame_sync's upload half).
 */
int lgame_set_state(lgame_handle *h, int slot, const lgame_state_view *state);

/*
 * coeff_1 * (state in slot_1) + coeff_2 * (state in slot_2), computed on the device and written
 * to a host state view: the densities, rhotheta_v, Exner pressure, winds and soil temperature
 * combine linearly; theta_v_pert is rhotheta_v / rho_gas - theta_v_bg of the combined state.
 * Replaces linear_combine_two_states (src/spatial_operators/mo_linear_combination.F90:15-37),
 * which the reference's time loop calls to interpolate the state to an output time
 * (control.F90:539).  slot_1 and slot_2 may be equal.
 */
int lgame_linear_combine(lgame_handle *h, int slot_1, int slot_2, double coeff_1, double coeff_2, lgame_state_view *out);

/*
 * Device -> host copy of resident slot 0 or 1 (the reference refreshes host state only before
 * write_output / write_out_integrals, control.F90:529-550).
 */
int lgame_get_state(lgame_handle *h, int slot, lgame_state_view *state);

/*
 * One HEVI predictor-corrector step on the resident slots: state_new = step(state_old).
 * Same arguments and meaning as pchevi (mo_manage_pchevi.F90:27): total_step_counter == 0
 * selects the first-step branches (mo_p_grad.F90:38-43,97-102; mo_vector_tend_expl.F90:58);
 * lrad_update asks for a radiation update before the substeps.  The carried diagnostics
 * (previous pressure gradient, lagged momentum advection, tendencies) live in the handle and
 * persist between calls exactly as t_diag/t_tend do between reference pchevi calls.
 * Asynchronous on the handle's stream; use lgame_sync to wait.
 */
int lgame_pchevi(lgame_handle *h, int slot_old, int slot_new, int total_step_counter,
                 int lrad_update, double t_0);

/*
 * Strict drop-in form: upload state_old and state_new, run one step, download state_new.
 * Mirrors the reference call signature with host arrays; costs two PCIe round trips.  Returns
 * LGAME_ENONFINITE when the step produced a non-finite value (as lgame_sync).
 */
int lgame_pchevi_host(lgame_handle *h, const lgame_state_view *state_old,
                      lgame_state_view *state_new, int total_step_counter, int lrad_update,
                      double t_0);

/*
 * Wait for all work queued on the handle's stream.  Returns LGAME_ENONFINITE if any step since
 * the last lgame_set_state produced a NaN or Inf in the new state: the corrector's column solvers
 * check every value they write (the densities, rhotheta_v, Exner pressure, theta_v and w), at no
 * extra memory traffic.  This is the reference's crash check (mo_write_out.F90:63-67, at every
 * output) made per step; the driver prints lgame_last_error and stops, as the reference does.
 */
int lgame_sync(lgame_handle *h);

/*
 * Domain budget sums of resident slot `slot` (write_out_integrals, mo_write_out.F90:226-304),
 * computed on the device in FP64 with compensated (two-sum) accumulation.  `out` receives
 * 4 + n_constituents doubles at the LGAME_INT_* slots.  The internal-energy sum uses the
 * handle's diagnosed temperature, as the reference does (control.F90:529-535 after pchevi).
 */
int lgame_integrals(lgame_handle *h, int slot, double *out);

/*
 * Finite check of resident slot `slot`: *n_bad = number of NaN/Inf values in the prognostic
 * fields (the reference's crash check, mo_write_out.F90:63-67, over every field).
 */
int lgame_check_finite(lgame_handle *h, int slot, long long *n_bad);

/* Copy a diagnostic field (LGAME_DIAG_*) to the host in the reference layout. */
int lgame_get_diag(lgame_handle *h, int which, double *out);

/*
 * Event timing of kernel groups on the handle's stream.  No counterpart in the reference, which
 * only prints a model speed (control.F90:544-547); for profiling tools and benchmarks.
 * lgame_profile(h, 1) starts recording (and clears earlier records); lgame_profile(h, 2) also
 * times every stencil kernel on its own (group LGAME_PROF_KERNEL0 + its index in the kernel order
 * of lgame_set_tiles; the scalar divergence kernel is index 14); lgame_profile_read waits for the
 * stream and returns the summed milliseconds and launch count of one group.
 */
#define LGAME_PROF_SOLVE_VER   0   /* three_band_solver_ver kernels                  */
#define LGAME_PROF_SOLVE_GEN   1   /* three_band_solver_gen_densities kernels        */
#define LGAME_PROF_STEP        2   /* whole lgame_pchevi calls                       */
#define LGAME_PROF_RADIATION   3   /* radiation updates (solver or surrogate + audit) */
#define LGAME_PROF_KERNEL0     8   /* + tile-table index: one stencil kernel (lgame_profile(h, 2)) */
int lgame_profile(lgame_handle *h, int enable);
int lgame_profile_read(lgame_handle *h, int group, double *total_ms, long long *count);
int lgame_profile_reset(lgame_handle *h);

/* ---- launch tuning ----
 * The horizontal stencil kernels run on blocks of 64 x-points by 1, 2 or 4 rows.  lgame_autotune
 * times every candidate for each of the LGAME_N_TILES kernels over `steps` steps of the handle's
 * own configuration and keeps the fastest; it must run after both slots are set and before the
 * first step, and restores the slots (and the zeroed work arrays) when it is done.  rows (may be
 * NULL) receives the choice, in the kernel order of lgame_set_tiles.  No reference counterpart:
 * the reference is a CPU code with OpenMP loops.
 * Kernel order: temperature, theta_fill, theta_grad, theta_div, theta_edge, pgrad, pot_vort,
 * mass_flux, vort_flux, ke, ke_grad, dissipation, vec_tend, scalar_flux, scalar_div, diff_zeta,
 * diff_coef, diff_dual, diff_mom, diff_flux, diff_div (the last six run only with diffusion).
 * The scalar fluxes are formed inside the scalar divergence kernel, the mass fluxes and the
 * kinetic energy inside the potential-vorticity kernel, the kinetic-energy gradient inside the
 * vorticity-flux kernel and the full theta_v of the third-order theta where it is loaded, so the
 * theta_fill, mass_flux, scalar_flux, ke and ke_grad entries are kept for the order but not used.  The pgrad entry is the first substep's
 * pressure gradient together with its wind update, vec_tend the second substep's wind update.
 * The scalar divergence kernel marches down the columns; its
 * levels can be cut into 1, 2, 4 or 8 segments (at least 4 levels each) so that small domains
 * fill the GPU: lgame_init picks the segments that give about four wavefronts per SIMD,
 * lgame_autotune times 1, 2, 4 and 8 after the block shapes, lgame_set_segments sets them and
 * lgame_get_segments reads the choice.  Every choice gives the same bits. */
#define LGAME_N_TILES 21
int lgame_autotune(lgame_handle *h, int steps, int *rows);
int lgame_set_tiles(lgame_handle *h, const int *rows);
int lgame_set_segments(lgame_handle *h, int nseg);
int lgame_get_segments(lgame_handle *h, int *nseg);
/* The implicit vertical solver (three_band_solver_ver, mo_column_solvers.F90:21-351) runs in one
 * pass per column (one lane a column) or in three passes: the rows of the system and the new
 * state by one thread per (column, level), the Thomas recursion alone per column.  lgame_init
 * picks three passes when the domain has fewer than 1024 wavefronts of 64 columns (fewer than the
 * GPU's SIMDs) and at least 3 layers, one pass otherwise; lgame_autotune times both on domains of
 * fewer than 2048 such wavefronts and keeps the faster; lgame_set_column_passes (1 or 3) sets it
 * and lgame_get_column_passes reads it.  Both give the same bits.  No reference counterpart. */
int lgame_set_column_passes(lgame_handle *h, int passes);
int lgame_get_column_passes(lgame_handle *h, int *passes);

/* ---- ensembles ----
 * Independent domains (handles, each with its own grid, size and state) advanced together.
 * mode 2 = batched: members whose steps run the same kernel sequence (same form, switches and
 * surrogate) form a group, and each kernel of the step is one launch over all members of the
 * group, with a work list of every member's blocks (radiation updates included: one gather, one
 * surrogate inference over all members' columns, one apply and one audit).  mode 0 = one launch
 * of a graph captured across all members' streams (one branch per member); steps at counter 0
 * and steps with a radiation update run eagerly.  mode 1 = the members one after another with
 * eager launches (the comparison).  Every mode gives the same bits.  The handles must stay alive
 * and unchanged in configuration, tiles and segments while the ensemble exists (the batched
 * groups are made at the first mode-2 step).  No reference counterpart (the reference runs one
 * domain per process). */
typedef struct lgame_ensemble lgame_ensemble;
int lgame_ensemble_create(lgame_handle *const *members, int n, lgame_ensemble **out);
int lgame_ensemble_step(lgame_ensemble *e, int total_step_counter, int lrad_update, double t_0, int mode);
int lgame_ensemble_sync(lgame_ensemble *e);
const char *lgame_ensemble_last_error(const lgame_ensemble *e);
void lgame_ensemble_destroy(lgame_ensemble *e);

/* ---- radiation boundary (update_rad_fluxes, src/radiation/mo_manage_radiation_calls.F90:17-59) ---- */

/* Per-column surface and location data; needed before the first step with lrad_update. */
int lgame_rad_set_surface(lgame_handle *h, const lgame_rad_view *surface);

/*
 * Inputs of the surrogate and outputs of the gray two-stream solver for every column of a state
 * slot at time coordinate t (the coszenith time): x (ny*nx, 5*n_layers+4) and y (ny*nx,
 * n_layers+2), row-major, columns in Fortran order ((ji-1) + ny*(jk-1)).  The temperature is
 * diagnosed from the slot first (it becomes the handle's diag temperature).  Either pointer may
 * be NULL.  This is how training data for the surrogate are harvested from the model's states.
 */
int lgame_rad_columns(lgame_handle *h, int slot, double t, float *x, float *y);

/* Radiation outputs of the last update: LGAME_RAD_OUT_TENDENCY (ny,nx,nl) [W m^-3],
   LGAME_RAD_OUT_SW_IN and LGAME_RAD_OUT_LW_OUT (ny,nx) [W m^-2], reference layout. */
#define LGAME_RAD_OUT_TENDENCY 0
#define LGAME_RAD_OUT_SW_IN    1
#define LGAME_RAD_OUT_LW_OUT   2
int lgame_rad_get(lgame_handle *h, int which, double *out);

typedef struct lgame_mlp lgame_mlp;

/*
 * Use a trained MLP (n_in = LGAME_RAD_N_IN, n_out = LGAME_RAD_N_OUT, with its normalisation set
 * by lgame_mlp_set_norm, else LGAME_EINVAL) as the radiation of a handle configured with
 * LGAME_RAD_SURROGATE.  The network works on LGAME_RAD_LEVELS levels: a domain of another layer
 * count is resampled, linearly in the layer index, to the network's levels and its heating rates
 * back to the domain's layers.  At every update a fraction audit_fraction of the columns (a
 * different sample each call, from seed) is recomputed with the solver; an audited column whose
 * heating-rate RMSE exceeds tol_k_per_day takes the solver's result, and so does every column whose
 * prediction leaves LGAME_RAD_ENVELOPE standard deviations of the training targets.  The handle
 * allocates its own inference buffers here (about 260 MB for 65,536 or more columns), so the step
 * allocates nothing and handles may share one network (weights are only read; training it while
 * handles use it is not allowed).  The handle keeps a pointer to m: m must outlive its use.
 */
int lgame_rad_attach_mlp(lgame_handle *h, lgame_mlp *m, double tol_k_per_day, double audit_fraction, unsigned long long seed);

/* Counters since lgame_init: radiation updates, columns of the random audit sample inside the
   training envelope, columns given back to the solver (over the tolerance or outside the
   envelope), and the RMSE [K/day] of the surrogate's heating rates over the counted sample. */
int lgame_rad_stats(lgame_handle *h, long long *calls, long long *audited, long long *fallbacks, double *audit_rmse);

/* ---- the surrogate network: a per-column MLP trained and evaluated with BF16 MFMA kernels ----
 * These entry points have no counterpart in the reference (which calls RTE+RRTMGP on every
 * radiation update); they serve the learned replacement of that call.
 * n_in -> hidden x n_hidden_layers (GELU) -> n_out; FP32 master weights with Adam, BF16
 * weights and activations in the GEMMs, FP32 accumulation.  Parameters are ordered W_1 [out][in],
 * b_1, W_2, b_2, ...  Inputs and targets are normalised with lgame_mlp_set_norm's mean and scale.
 */
int lgame_mlp_create(int n_in, int hidden, int n_hidden_layers, int n_out, unsigned long long seed, lgame_mlp **out);
long long lgame_mlp_n_params(const lgame_mlp *m);
int lgame_mlp_set_params(lgame_mlp *m, const float *params, long long n);
int lgame_mlp_get_params(lgame_mlp *m, float *params, long long n);
int lgame_mlp_set_norm(lgame_mlp *m, const float *x_mean, const float *x_scale, const float *y_mean, const float *y_scale);
/* y (n, n_out) = network(x (n, n_in)), host arrays, physical units */
int lgame_mlp_forward(lgame_mlp *m, const float *x, long long n, float *y);
/*
 * lgame_mlp_forward that allocates nothing, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments and results; the passes run in one static device
 * array of 8 MiB per GPU (held until the process ends), whose address is looked up once per GPU,
 * so a pass holds as many rows as fit (1,536 for the 304 -> 3x512 -> 62 network).  Returns
 * LGAME_EINVAL when 64 rows of the network do not fit (a hidden width above about 21,000 at three
 * hidden layers) or when another device than the network's is current.  Calls that overlap on one
 * GPU share the array and race: callers serialise them.
 */
int lgame_mlp_forward_noalloc(lgame_mlp *m, const float *x, long long n, float *y);
/* mean-squared error of the normalised outputs over n rows (a multiple of 64) and its gradient
   with respect to every parameter */
int lgame_mlp_gradient(lgame_mlp *m, const float *x, const float *y, long long n, float *grad, float *loss);
/* upload a training set (host arrays, physical units) */
int lgame_mlp_load_data(lgame_mlp *m, const float *x, const float *y, long long n);
/* Adam epochs over the loaded set in shuffled batches (a multiple of 64); loss[e] per epoch;
   *seconds the device time of the run */
int lgame_mlp_train(lgame_mlp *m, int epochs, int batch, float lr, unsigned long long seed, float *loss, double *seconds);
const char *lgame_mlp_last_error(const lgame_mlp *m);
void lgame_mlp_destroy(lgame_mlp *m);

/* Release the handle and its device memory. */
void lgame_finalize(lgame_handle *h);

/* Message of the last failing call on `h` (or of lgame_init when h is NULL). */
const char *lgame_last_error(const lgame_handle *h);

#ifdef __cplusplus
}
#endif

#endif /* LGAME_GPU_H */
