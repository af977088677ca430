! mo_lgame_gpu.F90 -- iso_c_binding interface of the MI355X HEVI step library (liblgame_gpu.so).
!
! The Fortran driver of L-GAME keeps its namelists, grid and state (t_grid, t_state of
! src/mo_definitions.F90) and calls the device step through this module instead of the
! reference pchevi (src/time_stepping/mo_manage_pchevi.F90:27).  The derived types below mirror
! the C structs of include/lgame_gpu.h field for field; logicals cross as integer(c_int).
! Arrays are passed with c_loc of the driver's allocatables, which are contiguous and keep the
! reference layout (y fastest).

module mo_lgame_gpu

  use iso_c_binding

  implicit none

  integer(c_int), parameter :: LGAME_OK = 0
  integer(c_int), parameter :: LGAME_PREC_FP64 = 0
  integer(c_int), parameter :: LGAME_PREC_MIXED = 1
  integer(c_int), parameter :: LGAME_RAD_OFF = 0

  type, bind(C) :: lgame_dims
    integer(c_int) :: ny, nx, n_layers, n_oro_layers, n_constituents, n_condensed_constituents, nsoillays
  end type lgame_dims

  type, bind(C) :: lgame_cfg
    real(c_double) :: dtime, toa, dx, dy, klemp_damp_max, klemp_begin_rel, dtime_rad, c_s, bg_shear, h_prandtl, karman
    integer(c_int) :: theta_adv_order, luse_bg_state, llinear, lcorio, lmoist, lklemp
    integer(c_int) :: lmom_diff_h, lmom_diff_v, ltemp_diff_h, ltemp_diff_v, lmass_diff_h, lmass_diff_v
    integer(c_int) :: lprog_soil_temp, lsfc_sensible_heat_flux, lsfc_phase_trans, lpbl
    integer(c_int) :: lperiodic, lfreeslip, lrad, precision, rad_mode, diff_coeff_scheme_h
  end type lgame_cfg

  type, bind(C) :: lgame_grid_view
    type(c_ptr) :: z_scalar, z_u, z_v, z_w, dx, dy, dz, area_x, area_y, area_z, volume, slope_x, slope_y
    type(c_ptr) :: inner_product_weights, area_dual_x, area_dual_y, area_dual_z, z_area_dual_z
    type(c_ptr) :: fvec_x, fvec_y, fvec_z, trsk_weights_u, trsk_weights_v
    type(c_ptr) :: exner_bg_grad_u, exner_bg_grad_v, exner_bg_grad_w, theta_v_bg, exner_bg
    type(c_ptr) :: gravity_m_v, gravity_potential, dx_dual, dy_dual, layer_thickness, land_fraction, roughness_length
    type(c_ptr) :: lake_fraction, sfc_rho_c, t_conduc_soil, t_const_soil, sst, z_soil_interface, z_soil_center
    real(c_double) :: mean_velocity_area, z_t_const
  end type lgame_grid_view

  ! lgame_grid_setup: generator parameters and the t_grid arrays it fills
  type, bind(C) :: lgame_grid_params
    real(c_double) :: dx, dy, toa, stretching_parameter, lat_center, lon_center
    integer(c_int) :: orography_id, lsleve, lplane, lfreeslip, n_hills
    type(c_ptr)    :: hills
    type(c_ptr)    :: oro  ! (ny,nx) orography with orography_id 1
  end type lgame_grid_params

  type, bind(C) :: lgame_grid_out
    type(c_ptr) :: z_scalar, z_u, z_v, z_w, dx, dy, dz, area_x, area_y, area_z, volume, slope_x, slope_y
    type(c_ptr) :: inner_product_weights, area_dual_x, area_dual_y, area_dual_z, z_area_dual_z
    type(c_ptr) :: fvec_x, fvec_y, fvec_z, trsk_weights_u, trsk_weights_v
    type(c_ptr) :: exner_bg_grad_u, exner_bg_grad_v, exner_bg_grad_w, theta_v_bg, exner_bg
    type(c_ptr) :: gravity_m_v, gravity_potential, dx_dual, dy_dual, layer_thickness, land_fraction, roughness_length
    type(c_ptr) :: lake_fraction, sfc_rho_c, t_conduc_soil, t_const_soil, sst, z_soil_interface, z_soil_center
    type(c_ptr) :: lat_geo_scalar, lon_geo_scalar
    real(c_double) :: mean_velocity_area, z_t_const
  end type lgame_grid_out

  type, bind(C) :: lgame_state_view
    type(c_ptr) :: rho, rhotheta_v, theta_v_pert, exner_pert, wind_u, wind_v, wind_w, temperature_soil
  end type lgame_state_view

  interface

    integer(c_int) function lgame_abi_version() bind(C,name="lgame_abi_version")
      import :: c_int
    end function lgame_abi_version

    ! grid_setup and bg_setup (mo_grid_generator.F90:29-1457) on the device
    integer(c_int) function lgame_grid_setup(dims,params,grid) bind(C,name="lgame_grid_setup")
      import :: c_int,lgame_dims,lgame_grid_params,lgame_grid_out
      type(lgame_dims),        intent(in)    :: dims
      type(lgame_grid_params), intent(in)    :: params
      type(lgame_grid_out),    intent(inout) :: grid
    end function lgame_grid_setup

    integer(c_int) function lgame_init(dims,cfg,grid,handle) bind(C,name="lgame_init")
      import :: c_int,c_ptr,lgame_dims,lgame_cfg,lgame_grid_view
      type(lgame_dims),      intent(in)  :: dims
      type(lgame_cfg),       intent(in)  :: cfg
      type(lgame_grid_view), intent(in)  :: grid
      type(c_ptr),           intent(out) :: handle
    end function lgame_init

    integer(c_int) function lgame_set_state(handle,slot,state) bind(C,name="lgame_set_state")
      import :: c_int,c_ptr,lgame_state_view
      type(c_ptr),            value,      intent(in) :: handle
      integer(c_int),         value,      intent(in) :: slot
      type(lgame_state_view),             intent(in) :: state
    end function lgame_set_state

    integer(c_int) function lgame_get_state(handle,slot,state) bind(C,name="lgame_get_state")
      import :: c_int,c_ptr,lgame_state_view
      type(c_ptr),            value,      intent(in)    :: handle
      integer(c_int),         value,      intent(in)    :: slot
      type(lgame_state_view),             intent(inout) :: state
    end function lgame_get_state

    ! state_write = coeff_1*slot_1 + coeff_2*slot_2 (linear_combine_two_states, control.F90:539)
    integer(c_int) function lgame_linear_combine(handle,slot_1,slot_2,coeff_1,coeff_2,state) &
                                                 bind(C,name="lgame_linear_combine")
      import :: c_int,c_ptr,c_double,lgame_state_view
      type(c_ptr),            value, intent(in)    :: handle
      integer(c_int),         value, intent(in)    :: slot_1,slot_2
      real(c_double),         value, intent(in)    :: coeff_1,coeff_2
      type(lgame_state_view),        intent(inout) :: state
    end function lgame_linear_combine

    integer(c_int) function lgame_pchevi(handle,slot_old,slot_new,total_step_counter,lrad_update,t_0) &
                                         bind(C,name="lgame_pchevi")
      import :: c_int,c_ptr,c_double
      type(c_ptr),    value, intent(in) :: handle
      integer(c_int), value, intent(in) :: slot_old,slot_new,total_step_counter,lrad_update
      real(c_double), value, intent(in) :: t_0
    end function lgame_pchevi

    integer(c_int) function lgame_sync(handle) bind(C,name="lgame_sync")
      import :: c_int,c_ptr
      type(c_ptr), value, intent(in) :: handle
    end function lgame_sync

    integer(c_int) function lgame_integrals(handle,slot,out) bind(C,name="lgame_integrals")
      import :: c_int,c_ptr,c_double
      type(c_ptr),    value, intent(in)    :: handle
      integer(c_int), value, intent(in)    :: slot
      real(c_double),        intent(inout) :: out(*)
    end function lgame_integrals

    integer(c_int) function lgame_check_finite(handle,slot,n_bad) bind(C,name="lgame_check_finite")
      import :: c_int,c_ptr,c_long_long
      type(c_ptr),        value, intent(in)  :: handle
      integer(c_int),     value, intent(in)  :: slot
      integer(c_long_long),      intent(out) :: n_bad
    end function lgame_check_finite

    ! the checkpoint of the fields the step carries between steps (an exact restart)
    integer(c_int) function lgame_checkpoint_size(handle,nbytes) bind(C,name="lgame_checkpoint_size")
      import :: c_int,c_ptr,c_long_long
      type(c_ptr),          value, intent(in)  :: handle
      integer(c_long_long),        intent(out) :: nbytes
    end function lgame_checkpoint_size

    integer(c_int) function lgame_checkpoint_save(handle,buf,nbytes) bind(C,name="lgame_checkpoint_save")
      import :: c_int,c_ptr,c_long_long
      type(c_ptr),          value, intent(in) :: handle
      type(c_ptr),          value, intent(in) :: buf
      integer(c_long_long), value, intent(in) :: nbytes
    end function lgame_checkpoint_save

    integer(c_int) function lgame_checkpoint_load(handle,buf,nbytes) bind(C,name="lgame_checkpoint_load")
      import :: c_int,c_ptr,c_long_long
      type(c_ptr),          value, intent(in) :: handle
      type(c_ptr),          value, intent(in) :: buf
      integer(c_long_long), value, intent(in) :: nbytes
    end function lgame_checkpoint_load

    subroutine lgame_finalize(handle) bind(C,name="lgame_finalize")
      import :: c_ptr
      type(c_ptr), value, intent(in) :: handle
    end subroutine lgame_finalize

    type(c_ptr) function lgame_last_error(handle) bind(C,name="lgame_last_error")
      import :: c_ptr
      type(c_ptr), value, intent(in) :: handle
    end function lgame_last_error

  end interface

contains

  ! Prints the library's last error message and stops like the reference does on fatal errors
  ! ("call exit(1)", mo_run_nml.F90:104-141).
  subroutine lgame_check(rc,handle,what)
    integer(c_int), intent(in) :: rc
    type(c_ptr),    intent(in) :: handle
    character(*),   intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc==LGAME_OK) return
    call c_f_pointer(lgame_last_error(handle),msg,[512])
    n = 0
    do while (n<512)
      if (msg(n+1)==c_null_char) exit
      n = n+1
    enddo
    write(*,*) "Error in ",what,": ",msg(1:n)
    call exit(1)
  end subroutine lgame_check

end module mo_lgame_gpu
