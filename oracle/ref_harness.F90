! ref_harness.F90 -- test harness around the reference's own compiled step modules.
!
! TEST INFRASTRUCTURE ONLY.  Built into oracle/_ref/libref.so (see oracle/Makefile) together with
! the unmodified reference modules that compile from their own sources: mo_definitions,
! mo_constants, the namelist modules, mo_dictionary, mo_derived, the spatial operators except
! mo_momentum_diff_diss, mo_tke, mo_eff_diff_coeffs, mo_pbl, mo_p_grad, mo_scalar_tend_expl and
! mo_column_solvers.  It exposes them to Python through bind(C) entry points so the C
! restatement in lgame_oracle.c can be checked against the reference arithmetic.
!
! mo_vector_tend_expl, mo_momentum_diff_diss and mo_manage_pchevi cannot be built here without
! stand-ins (they reach netCDF through mo_momentum_diff_diss -> mo_grid_generator, and RRTMGP
! through mo_manage_radiation_calls), so the glue they contribute is restated below:
!   * harness_vector_tend_expl follows mo_vector_tend_expl.F90:30-182 and calls the reference's
!     calc_pot_vort, scalar_times_vector_h/v, calc_vorticity_flux_term, inner_product, grad_vert,
!     grad_hor, update_n_squared and tke_update;
!   * harness_mom_diff_h / harness_mom_diff_v follow mom_diff_h / mom_diff_v,
!     mo_momentum_diff_diss.F90:25-331, and call the reference's div_h, rel_vort, hor_viscosity,
!     grad_hor, grad_vert, add_vertical_div and vert_vert_mom_viscosity;
!   * harness_dissipation follows simple_dissipation_rate, mo_momentum_diff_diss.F90:333-377;
!   * harness_pchevi follows pchevi, mo_manage_pchevi.F90:27-118.
! Everything else in a step (pressure gradient, scalar tendencies, both column solvers, the
! operators and thermodynamics) is the reference's own compiled code.

module ref_harness

  use iso_c_binding
  use mo_definitions,        only: wp,t_grid,t_state,t_diag,t_tend
  use mo_constants,          only: impl_thermo_weight,M_PI
  use mo_run_nml,            only: ny,nx,n_layers,n_levels,n_oro_layers,n_flat_layers,dy,dx,dtime,toa, &
                                   theta_adv_order,luse_bg_state,llinear,lcorio,lplane,eff_hor_res
  use mo_constituents_nml,   only: lmoist,n_condensed_constituents,n_gaseous_constituents,n_constituents
  use mo_diff_nml,           only: lmom_diff_h,lmom_diff_v,ltemp_diff_h,ltemp_diff_v,lmass_diff_h,lmass_diff_v, &
                                   lklemp,klemp_damp_max,klemp_begin_rel,diff_coeff_scheme_h,diff_coeff_scheme_v, &
                                   c_s,bg_shear,h_prandtl,karman
  use mo_surface_nml,        only: nsoillays,lprog_soil_temp,lsfc_sensible_heat_flux,lsfc_phase_trans,lpbl
  use mo_bc_nml,             only: lperiodic,lfreeslip
  use mo_derived,            only: temperature_diagnostics
  use mo_p_grad,             only: manage_p_grad,calc_p_grad_condensates_v
  use mo_scalar_tend_expl,   only: scalar_tend_expl
  use mo_column_solvers,     only: three_band_solver_ver,three_band_solver_gen_densities,thomas_algorithm
  use mo_inner_product,      only: inner_product,w_free_slip
  use mo_gradient_operators, only: grad_hor,grad_vert
  use mo_divergence_operators, only: div_h,add_vertical_div
  use mo_vorticities,        only: calc_pot_vort,rel_vort
  use mo_eff_diff_coeffs,    only: hor_viscosity,vert_vert_mom_viscosity,update_n_squared
  use mo_tke,                only: tke_update
  use mo_pbl,                only: update_sfc_turb_quantities,pbl_wind_tendency
  use mo_multiplications,    only: scalar_times_vector_h,scalar_times_vector_v
  use mo_vorticity_flux,     only: calc_vorticity_flux_term
  use mo_linear_combination, only: linear_combine_two_states

  implicit none

  type(t_grid),  save, target :: grid
  type(t_state), save, target :: st1,st2,st3
  type(t_tend),  save         :: tend
  type(t_diag),  save         :: diag
  integer,       save         :: n_damping_levels

contains

  subroutine alloc_state(s)
    type(t_state), intent(inout) :: s
    allocate(s%rho(ny,nx,n_layers,n_constituents)); s%rho = 0._wp
    allocate(s%rhotheta_v(ny,nx,n_layers)); s%rhotheta_v = 0._wp
    allocate(s%theta_v_pert(ny,nx,n_layers)); s%theta_v_pert = 0._wp
    allocate(s%exner_pert(ny,nx,n_layers)); s%exner_pert = 0._wp
    allocate(s%wind_u(ny,nx+1,n_layers)); s%wind_u = 0._wp
    allocate(s%wind_v(ny+1,nx,n_layers)); s%wind_v = 0._wp
    allocate(s%wind_w(ny,nx,n_levels)); s%wind_w = 0._wp
    allocate(s%temperature_soil(ny,nx,max(nsoillays,1))); s%temperature_soil = 0._wp
  end subroutine alloc_state

  ! idims: ny,nx,n_layers,n_oro_layers,n_constituents,n_condensed_constituents,nsoillays
  ! icfg:  theta_adv_order,luse_bg_state,llinear,lcorio,lmoist,lklemp,lperiodic,lfreeslip,
  !        lmom_diff_h,lmom_diff_v,ltemp_diff_h,ltemp_diff_v,lmass_diff_h,lmass_diff_v,
  !        diff_coeff_scheme_h (0 smag, 1 tke),lpbl,lprog_soil_temp,lsfc_sensible_heat_flux,lsfc_phase_trans
  ! rcfg:  dtime,toa,dx,dy,klemp_damp_max,klemp_begin_rel,c_s,bg_shear,mean_velocity_area,h_prandtl,karman
  subroutine ref_init(idims,icfg,rcfg) bind(C,name="ref_init")
    integer(c_int), intent(in) :: idims(7),icfg(19)
    real(c_double), intent(in) :: rcfg(11)

    ny = idims(1); nx = idims(2); n_layers = idims(3); n_oro_layers = idims(4)
    n_constituents = idims(5); n_condensed_constituents = idims(6); nsoillays = idims(7)
    n_levels = n_layers+1
    n_flat_layers = n_layers-n_oro_layers
    n_gaseous_constituents = n_constituents-n_condensed_constituents
    theta_adv_order = icfg(1)
    luse_bg_state = icfg(2)/=0
    llinear = icfg(3)/=0
    lcorio = icfg(4)/=0
    lmoist = icfg(5)/=0
    lklemp = icfg(6)/=0
    lperiodic = icfg(7)/=0
    lfreeslip = icfg(8)/=0
    dtime = rcfg(1); toa = rcfg(2); dx = rcfg(3); dy = rcfg(4)
    klemp_damp_max = rcfg(5); klemp_begin_rel = rcfg(6)
    lplane = .false.
    lmom_diff_h = icfg(9)/=0; lmom_diff_v = icfg(10)/=0; ltemp_diff_h = icfg(11)/=0; ltemp_diff_v = icfg(12)/=0
    lmass_diff_h = icfg(13)/=0; lmass_diff_v = icfg(14)/=0
    diff_coeff_scheme_h = "smag"
    if (icfg(15)==1) diff_coeff_scheme_h = "tke"
    diff_coeff_scheme_v = "tke"
    c_s = rcfg(7); bg_shear = rcfg(8); h_prandtl = rcfg(10); karman = rcfg(11)
    eff_hor_res = sqrt(dx*dy)
    lpbl = icfg(16)/=0; lprog_soil_temp = icfg(17)/=0; lsfc_sensible_heat_flux = icfg(18)/=0
    lsfc_phase_trans = icfg(19)/=0

    ! grid (sizes as in control.F90:101-157)
    allocate(grid%z_scalar(ny,nx,n_layers)); allocate(grid%z_u(ny,nx+1,n_layers))
    allocate(grid%z_v(ny+1,nx,n_layers)); allocate(grid%z_w(ny,nx,n_levels))
    allocate(grid%dx(ny,nx+1,n_layers)); allocate(grid%dy(ny+1,nx,n_layers)); allocate(grid%dz(ny,nx,n_levels))
    allocate(grid%area_x(ny,nx+1,n_layers)); allocate(grid%area_y(ny+1,nx,n_layers)); allocate(grid%area_z(ny,nx,n_levels))
    allocate(grid%volume(ny,nx,n_layers)); allocate(grid%slope_x(ny,nx+1,n_layers)); allocate(grid%slope_y(ny+1,nx,n_layers))
    allocate(grid%inner_product_weights(6,ny,nx,n_layers))
    allocate(grid%area_dual_x(ny+1,nx,n_levels)); allocate(grid%area_dual_y(ny,nx+1,n_levels))
    allocate(grid%area_dual_z(ny+1,nx+1,n_layers)); allocate(grid%z_area_dual_z(ny+1,nx+1,n_layers))
    allocate(grid%fvec_x(ny+1,nx)); allocate(grid%fvec_y(ny,nx+1)); allocate(grid%fvec_z(ny+1,nx+1))
    allocate(grid%trsk_weights_u(3,ny)); allocate(grid%trsk_weights_v(2,ny+1))
    allocate(grid%exner_bg_grad_u(ny,nx+1,n_layers)); allocate(grid%exner_bg_grad_v(ny+1,nx,n_layers))
    allocate(grid%exner_bg_grad_w(ny,nx,n_levels)); allocate(grid%theta_v_bg(ny,nx,n_layers))
    allocate(grid%exner_bg(ny,nx,n_layers)); allocate(grid%gravity_m_v(ny,nx,n_levels))
    allocate(grid%gravity_potential(ny,nx,n_layers))
    allocate(grid%dx_dual(ny+1,nx,n_layers)); allocate(grid%dy_dual(ny,nx+1,n_layers))
    allocate(grid%layer_thickness(ny,nx,n_layers))
    grid%mean_velocity_area = rcfg(9)
    allocate(grid%land_fraction(ny,nx)); grid%land_fraction = 0._wp
    allocate(grid%roughness_length(ny,nx)); grid%roughness_length = 0._wp
    allocate(grid%sfc_rho_c(ny,nx)); grid%sfc_rho_c = 0._wp
    allocate(grid%t_conduc_soil(ny,nx)); grid%t_conduc_soil = 0._wp
    allocate(grid%t_const_soil(ny,nx)); grid%t_const_soil = 0._wp
    allocate(grid%z_soil_interface(max(nsoillays,1)+1)); grid%z_soil_interface = 0._wp
    allocate(grid%z_soil_center(max(nsoillays,1))); grid%z_soil_center = 0._wp
    grid%z_t_const = 0._wp
    allocate(grid%lake_fraction(ny,nx)); grid%lake_fraction = 0._wp

    call alloc_state(st1)
    call alloc_state(st2)
    call alloc_state(st3)

    allocate(tend%rho(ny,nx,n_layers,n_constituents)); tend%rho = 0._wp
    allocate(tend%rhotheta_v(ny,nx,n_layers)); tend%rhotheta_v = 0._wp
    allocate(tend%wind_u(ny,nx+1,n_layers)); tend%wind_u = 0._wp
    allocate(tend%wind_v(ny+1,nx,n_layers)); tend%wind_v = 0._wp
    allocate(tend%wind_w(ny,nx,n_levels)); tend%wind_w = 0._wp

    ! the t_diag components the step touches (sizes as in control.F90:199-272), zeroed as at :370-443
    allocate(diag%v_squared(ny,nx,n_layers)); diag%v_squared = 0._wp
    allocate(diag%v_squared_grad_x(ny,nx+1,n_layers)); diag%v_squared_grad_x = 0._wp
    allocate(diag%v_squared_grad_y(ny+1,nx,n_layers)); diag%v_squared_grad_y = 0._wp
    allocate(diag%v_squared_grad_z(ny,nx,n_levels)); diag%v_squared_grad_z = 0._wp
    allocate(diag%p_grad_acc_neg_l_u(ny,nx+1,n_layers)); diag%p_grad_acc_neg_l_u = 0._wp
    allocate(diag%p_grad_acc_neg_l_v(ny+1,nx,n_layers)); diag%p_grad_acc_neg_l_v = 0._wp
    allocate(diag%p_grad_acc_neg_l_w(ny,nx,n_levels)); diag%p_grad_acc_neg_l_w = 0._wp
    allocate(diag%p_grad_acc_neg_nl_u(ny,nx+1,n_layers)); diag%p_grad_acc_neg_nl_u = 0._wp
    allocate(diag%p_grad_acc_neg_nl_v(ny+1,nx,n_layers)); diag%p_grad_acc_neg_nl_v = 0._wp
    allocate(diag%p_grad_acc_neg_nl_w(ny,nx,n_levels)); diag%p_grad_acc_neg_nl_w = 0._wp
    allocate(diag%p_grad_acc_old_u(ny,nx+1,n_layers)); diag%p_grad_acc_old_u = 0._wp
    allocate(diag%p_grad_acc_old_v(ny+1,nx,n_layers)); diag%p_grad_acc_old_v = 0._wp
    allocate(diag%p_grad_condensates_w(ny,nx,n_levels)); diag%p_grad_condensates_w = 0._wp
    allocate(diag%pot_vort_tend_x(ny,nx+1,n_layers)); diag%pot_vort_tend_x = 0._wp
    allocate(diag%pot_vort_tend_y(ny+1,nx,n_layers)); diag%pot_vort_tend_y = 0._wp
    allocate(diag%pot_vort_tend_z(ny,nx,n_levels)); diag%pot_vort_tend_z = 0._wp
    allocate(diag%scalar_placeholder(ny,nx,n_layers)); diag%scalar_placeholder = 0._wp
    allocate(diag%temperature(ny,nx,n_layers)); diag%temperature = 0._wp
    allocate(diag%u_placeholder(ny,nx+1,n_layers)); diag%u_placeholder = 0._wp
    allocate(diag%v_placeholder(ny+1,nx,n_layers)); diag%v_placeholder = 0._wp
    allocate(diag%w_placeholder(ny,nx,n_levels)); diag%w_placeholder = 0._wp
    allocate(diag%theta_v_u(ny,nx+1,n_layers)); diag%theta_v_u = 0._wp
    allocate(diag%theta_v_v(ny+1,nx,n_layers)); diag%theta_v_v = 0._wp
    allocate(diag%zeta_x(ny+1,nx,n_levels)); diag%zeta_x = 0._wp
    allocate(diag%zeta_y(ny,nx+1,n_levels)); diag%zeta_y = 0._wp
    allocate(diag%zeta_z(ny+1,nx+1,n_layers)); diag%zeta_z = 0._wp
    allocate(diag%eta_x(ny+1,nx,n_levels)); diag%eta_x = 0._wp
    allocate(diag%eta_y(ny,nx+1,n_levels)); diag%eta_y = 0._wp
    allocate(diag%eta_z(ny+1,nx+1,n_layers)); diag%eta_z = 0._wp
    allocate(diag%radiation_tendency(ny,nx,n_layers)); diag%radiation_tendency = 0._wp
    allocate(diag%flux_density_u(ny,nx+1,n_layers)); diag%flux_density_u = 0._wp
    allocate(diag%flux_density_v(ny+1,nx,n_layers)); diag%flux_density_v = 0._wp
    allocate(diag%flux_density_w(ny,nx,n_levels)); diag%flux_density_w = 0._wp
    allocate(diag%flux_density_div(ny,nx,n_layers)); diag%flux_density_div = 0._wp
    allocate(diag%p_grad_decel_factor(ny,nx,n_layers)); diag%p_grad_decel_factor = 0._wp
    allocate(diag%mom_diff_tend_x(ny,nx+1,n_layers)); diag%mom_diff_tend_x = 0._wp
    allocate(diag%mom_diff_tend_y(ny+1,nx,n_layers)); diag%mom_diff_tend_y = 0._wp
    allocate(diag%mom_diff_tend_z(ny,nx,n_levels)); diag%mom_diff_tend_z = 0._wp
    allocate(diag%heating_diss(ny,nx,n_layers)); diag%heating_diss = 0._wp
    allocate(diag%phase_trans_rates(ny,nx,n_layers,n_condensed_constituents+1)); diag%phase_trans_rates = 0._wp
    allocate(diag%phase_trans_heating_rate(ny,nx,n_layers)); diag%phase_trans_heating_rate = 0._wp
    allocate(diag%temp_diff_heating(ny,nx,n_layers)); diag%temp_diff_heating = 0._wp
    allocate(diag%condensates_sediment_heat(ny,nx,n_layers)); diag%condensates_sediment_heat = 0._wp
    allocate(diag%mass_diff_tendency(ny,nx,n_layers,n_constituents)); diag%mass_diff_tendency = 0._wp
    allocate(diag%a_rain(ny,nx,n_layers)); diag%a_rain = 0._wp
    allocate(diag%du_dz(ny,nx+1,n_levels)); diag%du_dz = 0._wp
    allocate(diag%dv_dz(ny+1,nx,n_levels)); diag%dv_dz = 0._wp
    allocate(diag%n_squared(ny,nx,n_layers)); diag%n_squared = 0._wp
    allocate(diag%tke(ny,nx,n_layers)); diag%tke = 0._wp
    allocate(diag%viscosity_molecular(ny,nx,n_layers)); diag%viscosity_molecular = 0._wp
    allocate(diag%viscosity(ny,nx,n_layers)); diag%viscosity = 0._wp
    allocate(diag%viscosity_dual(ny+1,nx+1,n_layers)); diag%viscosity_dual = 0._wp
    allocate(diag%vert_hor_viscosity_u(ny,nx+1,n_levels)); diag%vert_hor_viscosity_u = 0._wp
    allocate(diag%vert_hor_viscosity_v(ny+1,nx,n_levels)); diag%vert_hor_viscosity_v = 0._wp
    allocate(diag%mass_diff_coeff_eff_h(ny,nx,n_layers)); diag%mass_diff_coeff_eff_h = 0._wp
    allocate(diag%mass_diff_coeff_eff_v(ny,nx,n_layers)); diag%mass_diff_coeff_eff_v = 0._wp
    allocate(diag%temp_diff_coeff_eff_h(ny,nx,n_layers)); diag%temp_diff_coeff_eff_h = 0._wp
    allocate(diag%temp_diff_coeff_eff_v(ny,nx,n_layers)); diag%temp_diff_coeff_eff_v = 0._wp
    allocate(diag%roughness_length(ny,nx)); diag%roughness_length = 0._wp
    allocate(diag%roughness_velocity(ny,nx)); diag%roughness_velocity = 0._wp
    allocate(diag%monin_obukhov_length(ny,nx)); diag%monin_obukhov_length = 0._wp
    allocate(diag%scalar_flux_resistance(ny,nx)); diag%scalar_flux_resistance = 0._wp
    allocate(diag%sst(ny,nx)); diag%sst = 0._wp
    allocate(diag%power_flux_density_sens_soil(ny,nx)); diag%power_flux_density_sens_soil = 0._wp
    allocate(diag%power_flux_density_sens_sea(ny,nx)); diag%power_flux_density_sens_sea = 0._wp
    allocate(diag%power_flux_density_lat_lake(ny,nx)); diag%power_flux_density_lat_lake = 0._wp
    allocate(diag%power_flux_density_lat_sea(ny,nx)); diag%power_flux_density_lat_sea = 0._wp
    allocate(diag%sfc_sw_in(ny,nx)); diag%sfc_sw_in = 0._wp
    allocate(diag%sfc_lw_out(ny,nx)); diag%sfc_lw_out = 0._wp
  end subroutine ref_init

  ! grid field ids (Python side keeps the same table)
  subroutine ref_set_grid(id,src) bind(C,name="ref_set_grid")
    integer(c_int), value, intent(in) :: id
    type(c_ptr),    value, intent(in) :: src
    real(c_double), pointer :: p(:)
    integer :: n
    select case (id)
      case(1);  n = size(grid%z_scalar);      call c_f_pointer(src,p,[n]); grid%z_scalar = reshape(p,shape(grid%z_scalar))
      case(2);  n = size(grid%z_u);           call c_f_pointer(src,p,[n]); grid%z_u = reshape(p,shape(grid%z_u))
      case(3);  n = size(grid%z_v);           call c_f_pointer(src,p,[n]); grid%z_v = reshape(p,shape(grid%z_v))
      case(4);  n = size(grid%z_w);           call c_f_pointer(src,p,[n]); grid%z_w = reshape(p,shape(grid%z_w))
      case(5);  n = size(grid%dx);            call c_f_pointer(src,p,[n]); grid%dx = reshape(p,shape(grid%dx))
      case(6);  n = size(grid%dy);            call c_f_pointer(src,p,[n]); grid%dy = reshape(p,shape(grid%dy))
      case(7);  n = size(grid%dz);            call c_f_pointer(src,p,[n]); grid%dz = reshape(p,shape(grid%dz))
      case(8);  n = size(grid%area_x);        call c_f_pointer(src,p,[n]); grid%area_x = reshape(p,shape(grid%area_x))
      case(9);  n = size(grid%area_y);        call c_f_pointer(src,p,[n]); grid%area_y = reshape(p,shape(grid%area_y))
      case(10); n = size(grid%area_z);        call c_f_pointer(src,p,[n]); grid%area_z = reshape(p,shape(grid%area_z))
      case(11); n = size(grid%volume);        call c_f_pointer(src,p,[n]); grid%volume = reshape(p,shape(grid%volume))
      case(12); n = size(grid%slope_x);       call c_f_pointer(src,p,[n]); grid%slope_x = reshape(p,shape(grid%slope_x))
      case(13); n = size(grid%slope_y);       call c_f_pointer(src,p,[n]); grid%slope_y = reshape(p,shape(grid%slope_y))
      case(14); n = size(grid%inner_product_weights); call c_f_pointer(src,p,[n])
                grid%inner_product_weights = reshape(p,shape(grid%inner_product_weights))
      case(15); n = size(grid%area_dual_x);   call c_f_pointer(src,p,[n]); grid%area_dual_x = reshape(p,shape(grid%area_dual_x))
      case(16); n = size(grid%area_dual_y);   call c_f_pointer(src,p,[n]); grid%area_dual_y = reshape(p,shape(grid%area_dual_y))
      case(17); n = size(grid%area_dual_z);   call c_f_pointer(src,p,[n]); grid%area_dual_z = reshape(p,shape(grid%area_dual_z))
      case(18); n = size(grid%z_area_dual_z); call c_f_pointer(src,p,[n]); grid%z_area_dual_z = reshape(p,shape(grid%z_area_dual_z))
      case(19); n = size(grid%fvec_x);        call c_f_pointer(src,p,[n]); grid%fvec_x = reshape(p,shape(grid%fvec_x))
      case(20); n = size(grid%fvec_y);        call c_f_pointer(src,p,[n]); grid%fvec_y = reshape(p,shape(grid%fvec_y))
      case(21); n = size(grid%fvec_z);        call c_f_pointer(src,p,[n]); grid%fvec_z = reshape(p,shape(grid%fvec_z))
      case(22); n = size(grid%trsk_weights_u); call c_f_pointer(src,p,[n]); grid%trsk_weights_u = reshape(p,shape(grid%trsk_weights_u))
      case(23); n = size(grid%trsk_weights_v); call c_f_pointer(src,p,[n]); grid%trsk_weights_v = reshape(p,shape(grid%trsk_weights_v))
      case(24); n = size(grid%exner_bg_grad_u); call c_f_pointer(src,p,[n]); grid%exner_bg_grad_u = reshape(p,shape(grid%exner_bg_grad_u))
      case(25); n = size(grid%exner_bg_grad_v); call c_f_pointer(src,p,[n]); grid%exner_bg_grad_v = reshape(p,shape(grid%exner_bg_grad_v))
      case(26); n = size(grid%exner_bg_grad_w); call c_f_pointer(src,p,[n]); grid%exner_bg_grad_w = reshape(p,shape(grid%exner_bg_grad_w))
      case(27); n = size(grid%theta_v_bg);    call c_f_pointer(src,p,[n]); grid%theta_v_bg = reshape(p,shape(grid%theta_v_bg))
      case(28); n = size(grid%exner_bg);      call c_f_pointer(src,p,[n]); grid%exner_bg = reshape(p,shape(grid%exner_bg))
      case(29); n = size(grid%gravity_m_v);   call c_f_pointer(src,p,[n]); grid%gravity_m_v = reshape(p,shape(grid%gravity_m_v))
      case(30); n = size(grid%gravity_potential); call c_f_pointer(src,p,[n])
                grid%gravity_potential = reshape(p,shape(grid%gravity_potential))
      case(31); n = size(grid%dx_dual);       call c_f_pointer(src,p,[n]); grid%dx_dual = reshape(p,shape(grid%dx_dual))
      case(32); n = size(grid%dy_dual);       call c_f_pointer(src,p,[n]); grid%dy_dual = reshape(p,shape(grid%dy_dual))
      case(33); n = size(grid%layer_thickness); call c_f_pointer(src,p,[n])
                grid%layer_thickness = reshape(p,shape(grid%layer_thickness))
      case(34); n = size(grid%land_fraction); call c_f_pointer(src,p,[n]); grid%land_fraction = reshape(p,shape(grid%land_fraction))
      case(35); n = size(grid%roughness_length); call c_f_pointer(src,p,[n])
                grid%roughness_length = reshape(p,shape(grid%roughness_length))
      case(36); n = size(grid%lake_fraction); call c_f_pointer(src,p,[n]); grid%lake_fraction = reshape(p,shape(grid%lake_fraction))
      case(37); n = size(grid%sfc_rho_c); call c_f_pointer(src,p,[n]); grid%sfc_rho_c = reshape(p,shape(grid%sfc_rho_c))
      case(38); n = size(grid%t_conduc_soil); call c_f_pointer(src,p,[n]); grid%t_conduc_soil = reshape(p,shape(grid%t_conduc_soil))
      case(39); n = size(grid%t_const_soil); call c_f_pointer(src,p,[n]); grid%t_const_soil = reshape(p,shape(grid%t_const_soil))
      case(40); n = size(diag%sst); call c_f_pointer(src,p,[n]); diag%sst = reshape(p,shape(diag%sst))
      case(41); n = size(grid%z_soil_interface); call c_f_pointer(src,p,[n]); grid%z_soil_interface = p
      case(42); n = size(grid%z_soil_center); call c_f_pointer(src,p,[n]); grid%z_soil_center = p
      case(43); call c_f_pointer(src,p,[1]); grid%z_t_const = p(1)
    end select
  end subroutine ref_set_grid

  ! n_damping_levels, as mo_grid_generator.F90:1050-1058 derives it from z_w
  subroutine ref_finish_grid() bind(C,name="ref_finish_grid")
    integer :: jl
    ! initializing the diagnostic roughness length (control.F90:456-459)
    diag%roughness_length = grid%roughness_length
    n_damping_levels = 0
    do jl=1,n_levels
      if (maxval(grid%z_w(:,:,jl))>klemp_begin_rel*toa) then
        n_damping_levels = n_damping_levels+1
      endif
    enddo
  end subroutine ref_finish_grid

  function sel(slot) result(s)
    integer, intent(in) :: slot
    type(t_state), pointer :: s
    if (slot==0) then
      s => st1
    elseif (slot==1) then
      s => st2
    else
      s => st3
    endif
  end function sel

  ! state field ids: 1 rho, 2 rhotheta_v, 3 theta_v_pert, 4 exner_pert, 5 wind_u, 6 wind_v, 7 wind_w,
  ! 8 temperature_soil
  subroutine ref_state_io(slot,id,dir,buf) bind(C,name="ref_state_io")
    integer(c_int), value, intent(in) :: slot,id,dir
    type(c_ptr),    value, intent(in) :: buf
    type(t_state), pointer :: s
    real(c_double), pointer :: p(:)
    s => sel(slot)
    select case (id)
      case(1)
        call c_f_pointer(buf,p,[size(s%rho)])
        if (dir==0) then; s%rho = reshape(p,shape(s%rho)); else; p = reshape(s%rho,[size(s%rho)]); endif
      case(2)
        call c_f_pointer(buf,p,[size(s%rhotheta_v)])
        if (dir==0) then; s%rhotheta_v = reshape(p,shape(s%rhotheta_v)); else; p = reshape(s%rhotheta_v,[size(s%rhotheta_v)]); endif
      case(3)
        call c_f_pointer(buf,p,[size(s%theta_v_pert)])
        if (dir==0) then; s%theta_v_pert = reshape(p,shape(s%theta_v_pert)); else; p = reshape(s%theta_v_pert,[size(s%theta_v_pert)]); endif
      case(4)
        call c_f_pointer(buf,p,[size(s%exner_pert)])
        if (dir==0) then; s%exner_pert = reshape(p,shape(s%exner_pert)); else; p = reshape(s%exner_pert,[size(s%exner_pert)]); endif
      case(5)
        call c_f_pointer(buf,p,[size(s%wind_u)])
        if (dir==0) then; s%wind_u = reshape(p,shape(s%wind_u)); else; p = reshape(s%wind_u,[size(s%wind_u)]); endif
      case(6)
        call c_f_pointer(buf,p,[size(s%wind_v)])
        if (dir==0) then; s%wind_v = reshape(p,shape(s%wind_v)); else; p = reshape(s%wind_v,[size(s%wind_v)]); endif
      case(7)
        call c_f_pointer(buf,p,[size(s%wind_w)])
        if (dir==0) then; s%wind_w = reshape(p,shape(s%wind_w)); else; p = reshape(s%wind_w,[size(s%wind_w)]); endif
      case(8)
        call c_f_pointer(buf,p,[size(s%temperature_soil)])
        if (dir==0) then
          s%temperature_soil = reshape(p,shape(s%temperature_soil))
        else
          p = reshape(s%temperature_soil,[size(s%temperature_soil)])
        endif
    end select
  end subroutine ref_state_io

  ! diag field ids: 1 temperature, 2 heating_diss, 3 pot_vort_tend_x, 4 pot_vort_tend_z,
  ! 5 condensates_sediment_heat, 6 tke, 7 mom_diff_tend_x, 8 mom_diff_tend_y, 9 mom_diff_tend_z,
  ! 10 temp_diff_heating, 11 mass_diff_tendency, 12 viscosity, 13 n_squared, 14 roughness_length,
  ! 15 roughness_velocity, 16 monin_obukhov_length, 17 scalar_flux_resistance,
  ! 18 power_flux_density_sens_soil, 19 power_flux_density_sens_sea, 20-22 p_grad_acc_neg_nl_u/v/w,
  ! 23-25 p_grad_acc_neg_l_u/v/w, 26-27 p_grad_acc_old_u/v, 28 p_grad_condensates_w
  subroutine ref_get_diag(id,buf) bind(C,name="ref_get_diag")
    integer(c_int), value, intent(in) :: id
    type(c_ptr),    value, intent(in) :: buf
    real(c_double), pointer :: p(:)
    select case (id)
      case(1); call c_f_pointer(buf,p,[size(diag%temperature)]); p = reshape(diag%temperature,[size(diag%temperature)])
      case(2); call c_f_pointer(buf,p,[size(diag%heating_diss)]); p = reshape(diag%heating_diss,[size(diag%heating_diss)])
      case(3); call c_f_pointer(buf,p,[size(diag%pot_vort_tend_x)]); p = reshape(diag%pot_vort_tend_x,[size(diag%pot_vort_tend_x)])
      case(4); call c_f_pointer(buf,p,[size(diag%pot_vort_tend_z)]); p = reshape(diag%pot_vort_tend_z,[size(diag%pot_vort_tend_z)])
      case(5); call c_f_pointer(buf,p,[size(diag%condensates_sediment_heat)])
               p = reshape(diag%condensates_sediment_heat,[size(diag%condensates_sediment_heat)])
      case(6); call c_f_pointer(buf,p,[size(diag%tke)]); p = reshape(diag%tke,[size(diag%tke)])
      case(7); call c_f_pointer(buf,p,[size(diag%mom_diff_tend_x)]); p = reshape(diag%mom_diff_tend_x,[size(diag%mom_diff_tend_x)])
      case(8); call c_f_pointer(buf,p,[size(diag%mom_diff_tend_y)]); p = reshape(diag%mom_diff_tend_y,[size(diag%mom_diff_tend_y)])
      case(9); call c_f_pointer(buf,p,[size(diag%mom_diff_tend_z)]); p = reshape(diag%mom_diff_tend_z,[size(diag%mom_diff_tend_z)])
      case(10); call c_f_pointer(buf,p,[size(diag%temp_diff_heating)])
                p = reshape(diag%temp_diff_heating,[size(diag%temp_diff_heating)])
      case(11); call c_f_pointer(buf,p,[size(diag%mass_diff_tendency)])
                p = reshape(diag%mass_diff_tendency,[size(diag%mass_diff_tendency)])
      case(12); call c_f_pointer(buf,p,[size(diag%viscosity)]); p = reshape(diag%viscosity,[size(diag%viscosity)])
      case(13); call c_f_pointer(buf,p,[size(diag%n_squared)]); p = reshape(diag%n_squared,[size(diag%n_squared)])
      case(14); call c_f_pointer(buf,p,[size(diag%roughness_length)])
                p = reshape(diag%roughness_length,[size(diag%roughness_length)])
      case(15); call c_f_pointer(buf,p,[size(diag%roughness_velocity)])
                p = reshape(diag%roughness_velocity,[size(diag%roughness_velocity)])
      case(16); call c_f_pointer(buf,p,[size(diag%monin_obukhov_length)])
                p = reshape(diag%monin_obukhov_length,[size(diag%monin_obukhov_length)])
      case(17); call c_f_pointer(buf,p,[size(diag%scalar_flux_resistance)])
                p = reshape(diag%scalar_flux_resistance,[size(diag%scalar_flux_resistance)])
      case(18); call c_f_pointer(buf,p,[size(diag%power_flux_density_sens_soil)])
                p = reshape(diag%power_flux_density_sens_soil,[size(diag%power_flux_density_sens_soil)])
      case(19); call c_f_pointer(buf,p,[size(diag%power_flux_density_sens_sea)])
                p = reshape(diag%power_flux_density_sens_sea,[size(diag%power_flux_density_sens_sea)])
      case(20); call c_f_pointer(buf,p,[size(diag%p_grad_acc_neg_nl_u)])
                p = reshape(diag%p_grad_acc_neg_nl_u,[size(diag%p_grad_acc_neg_nl_u)])
      case(21); call c_f_pointer(buf,p,[size(diag%p_grad_acc_neg_nl_v)])
                p = reshape(diag%p_grad_acc_neg_nl_v,[size(diag%p_grad_acc_neg_nl_v)])
      case(22); call c_f_pointer(buf,p,[size(diag%p_grad_acc_neg_nl_w)])
                p = reshape(diag%p_grad_acc_neg_nl_w,[size(diag%p_grad_acc_neg_nl_w)])
      case(23); call c_f_pointer(buf,p,[size(diag%p_grad_acc_neg_l_u)])
                p = reshape(diag%p_grad_acc_neg_l_u,[size(diag%p_grad_acc_neg_l_u)])
      case(24); call c_f_pointer(buf,p,[size(diag%p_grad_acc_neg_l_v)])
                p = reshape(diag%p_grad_acc_neg_l_v,[size(diag%p_grad_acc_neg_l_v)])
      case(25); call c_f_pointer(buf,p,[size(diag%p_grad_acc_neg_l_w)])
                p = reshape(diag%p_grad_acc_neg_l_w,[size(diag%p_grad_acc_neg_l_w)])
      case(26); call c_f_pointer(buf,p,[size(diag%p_grad_acc_old_u)])
                p = reshape(diag%p_grad_acc_old_u,[size(diag%p_grad_acc_old_u)])
      case(27); call c_f_pointer(buf,p,[size(diag%p_grad_acc_old_v)])
                p = reshape(diag%p_grad_acc_old_v,[size(diag%p_grad_acc_old_v)])
      case(28); call c_f_pointer(buf,p,[size(diag%p_grad_condensates_w)])
                p = reshape(diag%p_grad_condensates_w,[size(diag%p_grad_condensates_w)])
    end select
  end subroutine ref_get_diag

  ! simple_dissipation_rate restated (mo_momentum_diff_diss.F90:333-377)
  subroutine harness_dissipation(state)
    type(t_state), intent(in) :: state
    integer  :: ji,jk,jl
    real(wp) :: damping_start_height,z_above_damping,damping_coeff,damping_prefactor
    diag%w_placeholder = diag%mom_diff_tend_z
    if (lklemp) then
      damping_start_height = klemp_begin_rel*toa
      do jl=2,n_damping_levels
        do jk=1,nx
          do ji=1,ny
            z_above_damping = grid%z_w(ji,jk,jl) - damping_start_height
            if (z_above_damping<0._wp) then
              damping_coeff = 0._wp
            else
              damping_coeff = klemp_damp_max*sin(0.5_wp*M_PI*z_above_damping/(toa - damping_start_height))**2
            endif
            damping_prefactor = 1._wp + dtime*damping_coeff
            diag%w_placeholder(ji,jk,jl) = diag%w_placeholder(ji,jk,jl) &
                                           + (state%wind_w(ji,jk,jl)/damping_prefactor - state%wind_w(ji,jk,jl))/dtime
          enddo
        enddo
      enddo
    endif
    call inner_product(state%wind_u,state%wind_v,state%wind_w,diag%mom_diff_tend_x,diag%mom_diff_tend_y, &
                       diag%w_placeholder,diag%heating_diss,grid)
    diag%heating_diss = -sum(state%rho(:,:,:,1:n_condensed_constituents+1),4)*diag%heating_diss
  end subroutine harness_dissipation

  ! the averaged density sum(rho(1:ncond+1)) of two cells, as mo_momentum_diff_diss.F90 forms it
  pure function mean_rho2(state,ia,ka,la,ib,kb,lb) result(r)
    type(t_state), intent(in) :: state
    integer,       intent(in) :: ia,ka,la,ib,kb,lb
    real(wp) :: r
    r = 0.5_wp*(sum(state%rho(ia,ka,la,1:n_condensed_constituents+1)) &
              + sum(state%rho(ib,kb,lb,1:n_condensed_constituents+1)))
  end function mean_rho2

  ! vertical derivative of viscosity_dual*zeta_z at dual point (ji,jk) around layer jl, halved
  ! (the two terms of the terrain-following correction, mo_momentum_diff_diss.F90:75-82)
  pure function half_dvz(ji,jk,jl) result(r)
    integer, intent(in) :: ji,jk,jl
    real(wp) :: r
    integer  :: lu,ll
    lu = max(1,jl-1)
    ll = min(n_layers,jl+1)
    r = 0.5_wp*(diag%viscosity_dual(ji,jk,lu)*diag%zeta_z(ji,jk,lu) - diag%viscosity_dual(ji,jk,ll)*diag%zeta_z(ji,jk,ll)) &
        /(grid%z_area_dual_z(ji,jk,lu) - grid%z_area_dual_z(ji,jk,ll))
  end function half_dvz

  ! mom_diff_h restated (mo_momentum_diff_diss.F90:25-160), periodic boundaries
  subroutine harness_mom_diff_h(state)
    type(t_state), intent(in) :: state
    integer  :: ji,jk,jl
    real(wp) :: slope
    call div_h(state%wind_u,state%wind_v,diag%scalar_placeholder,grid)
    call rel_vort(state,diag,grid)
    call hor_viscosity(state,diag,grid)
    diag%scalar_placeholder = diag%viscosity*diag%scalar_placeholder
    ! the terrain correction of grad_hor reads mom_diff_tend_z as it is at this point
    call grad_hor(diag%scalar_placeholder,diag%mom_diff_tend_x,diag%mom_diff_tend_y,diag%mom_diff_tend_z,grid)
    do jl=1,n_layers
      do jk=1,nx+1
        do ji=1,ny
          diag%u_placeholder(ji,jk,jl) = (diag%viscosity_dual(ji+1,jk,jl)*diag%zeta_z(ji+1,jk,jl) - &
                                          diag%viscosity_dual(ji,jk,jl)*diag%zeta_z(ji,jk,jl))/grid%dy_dual(ji,jk,jl)
          slope = (grid%z_area_dual_z(ji+1,jk,jl) - grid%z_area_dual_z(ji,jk,jl))/grid%dy_dual(ji,jk,jl)
          diag%u_placeholder(ji,jk,jl) = diag%u_placeholder(ji,jk,jl) - slope*(half_dvz(ji+1,jk,jl) + half_dvz(ji,jk,jl))
        enddo
      enddo
    enddo
    do jl=1,n_layers
      do jk=1,nx
        do ji=1,ny+1
          diag%v_placeholder(ji,jk,jl) = (diag%viscosity_dual(ji,jk+1,jl)*diag%zeta_z(ji,jk+1,jl) - &
                                          diag%viscosity_dual(ji,jk,jl)*diag%zeta_z(ji,jk,jl))/grid%dx_dual(ji,jk,jl)
          slope = (grid%z_area_dual_z(ji,jk+1,jl) - grid%z_area_dual_z(ji,jk,jl))/grid%dx_dual(ji,jk,jl)
          diag%v_placeholder(ji,jk,jl) = diag%v_placeholder(ji,jk,jl) - slope*(half_dvz(ji,jk+1,jl) + half_dvz(ji,jk,jl))
        enddo
      enddo
    enddo
    do jl=1,n_layers
      do ji=1,ny
        do jk=1,nx
          diag%mom_diff_tend_x(ji,jk,jl) = (diag%mom_diff_tend_x(ji,jk,jl) + diag%u_placeholder(ji,jk,jl)) &
                                           /mean_rho2(state,ji,merge(nx,jk-1,jk==1),jl,ji,jk,jl)
        enddo
        diag%mom_diff_tend_x(ji,nx+1,jl) = diag%mom_diff_tend_x(ji,1,jl)
      enddo
    enddo
    do jl=1,n_layers
      do jk=1,nx
        do ji=1,ny
          diag%mom_diff_tend_y(ji,jk,jl) = (diag%mom_diff_tend_y(ji,jk,jl) + diag%v_placeholder(ji,jk,jl)) &
                                           /mean_rho2(state,merge(ny,ji-1,ji==1),jk,jl,ji,jk,jl)
        enddo
        diag%mom_diff_tend_y(ny+1,jk,jl) = diag%mom_diff_tend_y(1,jk,jl)
      enddo
    enddo
  end subroutine harness_mom_diff_h

  ! mom_diff_v restated (mo_momentum_diff_diss.F90:162-331), periodic boundaries
  subroutine harness_mom_diff_v(state)
    type(t_state), intent(in) :: state
    integer :: ji,jk,jl,ka,ia
    ! vertical gradients of u and v at the levels
    do jl=2,n_layers
      diag%du_dz(:,:,jl) = (state%wind_u(:,:,jl-1) - state%wind_u(:,:,jl))/(grid%z_u(:,:,jl-1) - grid%z_u(:,:,jl))
      diag%dv_dz(:,:,jl) = (state%wind_v(:,:,jl-1) - state%wind_v(:,:,jl))/(grid%z_v(:,:,jl-1) - grid%z_v(:,:,jl))
    enddo
    diag%du_dz(:,:,1) = diag%du_dz(:,:,2)
    diag%dv_dz(:,:,1) = diag%dv_dz(:,:,2)
    do jk=1,nx
      ka = merge(nx,jk-1,jk==1)
      diag%du_dz(:,jk,n_levels) = state%wind_u(:,jk,n_layers)/(grid%z_u(:,jk,n_layers) &
                                  - 0.5_wp*(grid%z_w(:,ka,n_levels)+grid%z_w(:,jk,n_levels)))
    enddo
    diag%du_dz(:,nx+1,n_levels) = diag%du_dz(:,1,n_levels)
    do ji=1,ny
      ia = merge(ny,ji-1,ji==1)
      diag%dv_dz(ji,:,n_levels) = state%wind_v(ji,:,n_layers)/(grid%z_v(ji,:,n_layers) &
                                  - 0.5_wp*(grid%z_w(ia,:,n_levels)+grid%z_w(ji,:,n_levels)))
    enddo
    diag%dv_dz(ny+1,:,n_levels) = diag%dv_dz(1,:,n_levels)
    ! their vertical diffusion with vert_hor_viscosity_u/_v
    do jl=1,n_layers
      do ji=1,ny
        do jk=1,nx
          ka = merge(nx,jk-1,jk==1)
          diag%mom_diff_tend_x(ji,jk,jl) = diag%mom_diff_tend_x(ji,jk,jl) &
          + (diag%vert_hor_viscosity_u(ji,jk,jl)*diag%du_dz(ji,jk,jl) &
          - diag%vert_hor_viscosity_u(ji,jk,jl+1)*diag%du_dz(ji,jk,jl+1)) &
          /(0.5_wp*(grid%z_w(ji,ka,jl) + grid%z_w(ji,jk,jl) - grid%z_w(ji,ka,jl+1) - grid%z_w(ji,jk,jl+1))) &
          /mean_rho2(state,ji,ka,jl,ji,jk,jl)
        enddo
        diag%mom_diff_tend_x(ji,nx+1,jl) = diag%mom_diff_tend_x(ji,1,jl)
      enddo
    enddo
    do jl=1,n_layers
      do jk=1,nx
        do ji=1,ny
          ia = merge(ny,ji-1,ji==1)
          diag%mom_diff_tend_y(ji,jk,jl) = diag%mom_diff_tend_y(ji,jk,jl) &
          + (diag%vert_hor_viscosity_v(ji,jk,jl)*diag%dv_dz(ji,jk,jl) &
          - diag%vert_hor_viscosity_v(ji,jk,jl+1)*diag%dv_dz(ji,jk,jl+1)) &
          /(0.5_wp*(grid%z_w(ia,jk,jl) + grid%z_w(ji,jk,jl) - grid%z_w(ia,jk,jl+1) - grid%z_w(ji,jk,jl+1))) &
          /mean_rho2(state,ia,jk,jl,ji,jk,jl)
        enddo
        diag%mom_diff_tend_y(ny+1,jk,jl) = diag%mom_diff_tend_y(1,jk,jl)
      enddo
    enddo
    ! vertical diffusion of w
    diag%scalar_placeholder = 0._wp
    call add_vertical_div(state%wind_w,diag%scalar_placeholder,grid)
    call vert_vert_mom_viscosity(state,diag,grid)
    call grad_vert(diag%scalar_placeholder,diag%mom_diff_tend_z,grid)
    ! horizontal diffusion of w (w_placeholder is read as it stands, as in the reference)
    do jl=1,n_layers
      diag%scalar_placeholder(:,:,jl) = grid%inner_product_weights(5,:,:,jl)*state%wind_w(:,:,jl) &
                                        + grid%inner_product_weights(6,:,:,jl)*state%wind_w(:,:,jl+1)
    enddo
    call grad_hor(diag%scalar_placeholder,diag%u_placeholder,diag%v_placeholder,diag%w_placeholder,grid)
    do jl=1,n_layers
      diag%u_placeholder(:,:,jl) = 0.5_wp*(diag%vert_hor_viscosity_u(:,:,jl) + diag%vert_hor_viscosity_u(:,:,jl+1)) &
                                   *diag%u_placeholder(:,:,jl)
      diag%v_placeholder(:,:,jl) = 0.5_wp*(diag%vert_hor_viscosity_v(:,:,jl) + diag%vert_hor_viscosity_v(:,:,jl+1)) &
                                   *diag%v_placeholder(:,:,jl)
    enddo
    call div_h(diag%u_placeholder,diag%v_placeholder,diag%scalar_placeholder,grid)
    do jl=2,n_layers
      do jk=1,nx
        do ji=1,ny
          diag%mom_diff_tend_z(ji,jk,jl) = diag%mom_diff_tend_z(ji,jk,jl) &
                                           + 0.5_wp*(diag%scalar_placeholder(ji,jk,jl-1) + diag%scalar_placeholder(ji,jk,jl))
          diag%mom_diff_tend_z(ji,jk,jl) = diag%mom_diff_tend_z(ji,jk,jl)/mean_rho2(state,ji,jk,jl-1,ji,jk,jl)
        enddo
      enddo
    enddo
  end subroutine harness_mom_diff_v

  ! vector_tend_expl restated (mo_vector_tend_expl.F90:30-182)
  subroutine harness_vector_tend_expl(state,rk_step,total_step_counter)
    type(t_state), intent(in) :: state
    integer,       intent(in) :: rk_step,total_step_counter
    integer  :: ji,jk,no_bg_switch
    real(wp) :: density_value,old_hor_pgrad_weight,current_hor_pgrad_weight,current_ver_pgrad_weight
    real(wp) :: old_weight,new_weight

    no_bg_switch = 0
    if (.not. luse_bg_state) then
      no_bg_switch = 1
    endif
    if ((rk_step==2 .or. total_step_counter==0) .and. ((.not. llinear) .or. lcorio)) then
      call calc_pot_vort(state,diag,grid)
      call scalar_times_vector_h(state%rho(:,:,:,n_condensed_constituents+1),state%wind_u,state%wind_v, &
                                 diag%u_placeholder,diag%v_placeholder)
      call scalar_times_vector_v(state%rho(:,:,:,n_condensed_constituents+1),state%wind_w,diag%w_placeholder)
      if (lfreeslip) then
        do jk=1,nx
          do ji=1,ny
            density_value = state%rho(ji,jk,n_layers,n_condensed_constituents+1) &
            + (state%rho(ji,jk,n_layers-1,n_condensed_constituents+1)-state%rho(ji,jk,n_layers,n_condensed_constituents+1)) &
            /(grid%z_scalar(ji,jk,n_layers-1)-grid%z_scalar(ji,jk,n_layers)) &
            *(grid%z_w(ji,jk,n_levels)-grid%z_scalar(ji,jk,n_layers))
            diag%w_placeholder(ji,jk,n_levels) = density_value*state%wind_w(ji,jk,n_levels)
          enddo
        enddo
      endif
      call calc_vorticity_flux_term(diag,grid)
      if (lfreeslip) then
        diag%w_placeholder(:,:,n_levels) = 0._wp
      endif
      if (.not. llinear) then
        call inner_product(state%wind_u,state%wind_v,state%wind_w,state%wind_u,state%wind_v,state%wind_w,diag%v_squared,grid)
        call grad_vert(diag%v_squared,diag%v_squared_grad_z,grid)
        call grad_hor(diag%v_squared,diag%v_squared_grad_x,diag%v_squared_grad_y,diag%v_squared_grad_z,grid)
      endif
    endif
    if (rk_step==1) then
      if (lmom_diff_h .or. lmass_diff_h .or. ltemp_diff_h) then
        call update_n_squared(state,diag,grid)
        if (diff_coeff_scheme_h=="tke" .or. diff_coeff_scheme_v=="tke") then
          call tke_update(state,diag,grid)
        endif
      endif
      if (lmom_diff_h) then
        call harness_mom_diff_h(state)
      endif
      if (lmom_diff_v) then
        call harness_mom_diff_v(state)
      endif
      if (lpbl) then
        call pbl_wind_tendency(state,diag,grid)
      endif
      if (lmom_diff_h .or. lpbl .or. lklemp) then
        call harness_dissipation(state)
      endif
    endif
    new_weight = 1._wp
    if (rk_step==2) then
      new_weight = 0.5_wp
    endif
    old_weight = 1._wp - new_weight
    current_hor_pgrad_weight = 0.5_wp + impl_thermo_weight
    old_hor_pgrad_weight = 1._wp - current_hor_pgrad_weight
    current_ver_pgrad_weight = 1._wp - impl_thermo_weight
    tend%wind_u = old_weight*tend%wind_u &
    + new_weight*(old_hor_pgrad_weight*diag%p_grad_acc_old_u &
    - current_hor_pgrad_weight*(diag%p_grad_acc_neg_nl_u + diag%p_grad_acc_neg_l_u) &
    - 0.5_wp*diag%v_squared_grad_x + diag%pot_vort_tend_x + diag%mom_diff_tend_x)
    tend%wind_v = old_weight*tend%wind_v &
    + new_weight*(old_hor_pgrad_weight*diag%p_grad_acc_old_v &
    - current_hor_pgrad_weight*(diag%p_grad_acc_neg_nl_v + diag%p_grad_acc_neg_l_v) &
    - 0.5_wp*diag%v_squared_grad_y + diag%pot_vort_tend_y + diag%mom_diff_tend_y)
    tend%wind_w = old_weight*tend%wind_w &
    + new_weight*(- current_ver_pgrad_weight*(diag%p_grad_acc_neg_nl_w + diag%p_grad_acc_neg_l_w) &
    - no_bg_switch*(1._wp-current_ver_pgrad_weight)*diag%p_grad_acc_neg_l_w &
    - 0.5_wp*diag%v_squared_grad_z + diag%pot_vort_tend_z + diag%p_grad_condensates_w + diag%mom_diff_tend_z)
  end subroutine harness_vector_tend_expl

  ! pchevi restated (mo_manage_pchevi.F90:27-118): no radiation or boundary updates
  subroutine harness_pchevi(state_old,state_new,total_step_counter)
    type(t_state), intent(inout) :: state_old,state_new
    integer,       intent(in)    :: total_step_counter
    integer :: rk_step
    call temperature_diagnostics(state_old,diag,grid)
    if (lsfc_sensible_heat_flux .or. lsfc_phase_trans .or. lpbl) then
      call update_sfc_turb_quantities(state_old,diag,grid)
    endif
    do rk_step=1,2
      if (rk_step==1) then
        call manage_p_grad(state_old,diag,grid,total_step_counter==0)
      endif
      if (rk_step==1) then
        call calc_p_grad_condensates_v(state_old,diag,grid)
        call harness_vector_tend_expl(state_old,rk_step,total_step_counter)
      endif
      if (rk_step==2) then
        call calc_p_grad_condensates_v(state_old,diag,grid)
        call harness_vector_tend_expl(state_new,rk_step,total_step_counter)
      endif
      state_new%wind_u = state_old%wind_u + dtime*tend%wind_u
      state_new%wind_v = state_old%wind_v + dtime*tend%wind_v
      call scalar_tend_expl(grid,state_old,state_new,tend,diag,rk_step)
      call three_band_solver_ver(state_old,state_new,diag,tend,grid,rk_step)
      if (lfreeslip) then
        call w_free_slip(state_new,grid)
      endif
      if (n_constituents>1) then
        call three_band_solver_gen_densities(state_old,state_new,tend,diag,grid,rk_step)
      endif
    enddo
  end subroutine harness_pchevi

  ! nsteps steps with the reference's ping-pong (control.F90:520-524) starting at counter `first`
  subroutine ref_run(first,nsteps) bind(C,name="ref_run")
    integer(c_int), value, intent(in) :: first,nsteps
    integer :: s
    do s=first,first+nsteps-1
      if (mod(s,2)==0) then
        call harness_pchevi(st1,st2,s)
      else
        call harness_pchevi(st2,st1,s)
      endif
    enddo
  end subroutine ref_run

  ! slot 2 = coeff_1*(slot slot_1) + coeff_2*(slot slot_2), linear_combine_two_states
  ! (mo_linear_combination.F90:15-37), as control.F90:539 interpolates to an output time
  subroutine ref_linear_combine(slot_1,slot_2,coeff_1,coeff_2) bind(C,name="ref_linear_combine")
    integer(c_int), value, intent(in) :: slot_1,slot_2
    real(c_double), value, intent(in) :: coeff_1,coeff_2
    call linear_combine_two_states(sel(slot_1),sel(slot_2),st3,coeff_1,coeff_2,grid)
  end subroutine ref_linear_combine

  ! the reference's manage_p_grad (mo_p_grad.F90:20-104) and calc_p_grad_condensates_v
  ! (:106-120) on a state slot, with the diagnostics as they stand
  subroutine ref_manage_p_grad(slot,lfirst) bind(C,name="ref_manage_p_grad")
    integer(c_int), value, intent(in) :: slot,lfirst
    call manage_p_grad(sel(slot),diag,grid,lfirst/=0)
    call calc_p_grad_condensates_v(sel(slot),diag,grid)
  end subroutine ref_manage_p_grad

  ! state_2 = state_1 (control.F90:500)
  subroutine ref_copy_slot0_to_1() bind(C,name="ref_copy_slot0_to_1")
    st2 = st1
  end subroutine ref_copy_slot0_to_1

  ! Thomas algorithm of the reference (mo_column_solvers.F90:606-640) on caller vectors
  subroutine ref_thomas(c,d,e,r,x,n) bind(C,name="ref_thomas")
    integer(c_int), value, intent(in) :: n
    real(c_double), intent(in)    :: c(n-1),d(n),e(n-1),r(n)
    real(c_double), intent(inout) :: x(n)
    call thomas_algorithm(c,d,e,r,x,n)
  end subroutine ref_thomas

end module ref_harness
