! lgame_driver_demo.F90 -- minimal Fortran time loop driving the GPU step through mo_lgame_gpu.
!
! It mirrors the parts of src/control.F90 around pchevi: allocate the grid and the two states
! with the reference extents, set them up (here read from a raw file written by the caller),
! copy state 1 to state 2 (control.F90:500), run the ping-pong loop (control.F90:507-556) and write
! the final state.  Usage:  lgame_driver_demo <case.bin> <out.bin> <nsteps>
! case.bin: int32 header [ny nx n_layers n_oro_layers n_constituents n_condensed nsoillays
! theta_adv_order lcorio lmoist lklemp lfreeslip llinear], float64 [dtime toa dx dy klemp_damp_max
! klemp_begin_rel], then the 30 grid arrays and the 7 state arrays in the order of
! include/lgame_gpu.h, each in Fortran order.

program lgame_driver_demo

  use iso_c_binding
  use mo_lgame_gpu

  implicit none

  integer(c_int)         :: hdr(13)
  real(c_double)         :: rc_(6)
  type(lgame_dims)       :: dims
  type(lgame_cfg)        :: cfg
  type(lgame_grid_view)  :: gv
  type(lgame_state_view) :: sv1,sv2
  type(c_ptr)            :: h
  integer                :: ny,nx,nl,nc,nsteps,step,u,rc
  character(len=512)     :: fin,fout,arg
  real(c_double), allocatable, target :: z_scalar(:,:,:),z_u(:,:,:),z_v(:,:,:),z_w(:,:,:),dx(:,:,:),dy(:,:,:),dz(:,:,:)
  real(c_double), allocatable, target :: area_x(:,:,:),area_y(:,:,:),area_z(:,:,:),volume(:,:,:),slope_x(:,:,:),slope_y(:,:,:)
  real(c_double), allocatable, target :: ipw(:,:,:,:),area_dual_x(:,:,:),area_dual_y(:,:,:),area_dual_z(:,:,:),z_area_dual_z(:,:,:)
  real(c_double), allocatable, target :: fvec_x(:,:),fvec_y(:,:),fvec_z(:,:),trsk_u(:,:),trsk_v(:,:)
  real(c_double), allocatable, target :: ebg_u(:,:,:),ebg_v(:,:,:),ebg_w(:,:,:),theta_v_bg(:,:,:),exner_bg(:,:,:)
  real(c_double), allocatable, target :: gravity_m_v(:,:,:),gravity_potential(:,:,:)
  real(c_double), allocatable, target :: rho1(:,:,:,:),rt1(:,:,:),tp1(:,:,:),ep1(:,:,:),u1(:,:,:),v1(:,:,:),w1(:,:,:)
  real(c_double), allocatable, target :: rho2(:,:,:,:),rt2(:,:,:),tp2(:,:,:),ep2(:,:,:),u2(:,:,:),v2(:,:,:),w2(:,:,:)

  call get_command_argument(1,fin)
  call get_command_argument(2,fout)
  call get_command_argument(3,arg)
  read(arg,*) nsteps

  open(newunit=u,file=trim(fin),access="stream",form="unformatted",status="old")
  read(u) hdr
  read(u) rc_
  ny = hdr(1); nx = hdr(2); nl = hdr(3); nc = hdr(5)
  dims = lgame_dims(hdr(1),hdr(2),hdr(3),hdr(4),hdr(5),hdr(6),hdr(7))
  cfg = lgame_cfg(rc_(1),rc_(2),rc_(3),rc_(4),rc_(5),rc_(6),60._c_double*rc_(4)/1000._c_double,0.2_c_double,1.5e-5_c_double,100._c_double,0.4_c_double, &
                  hdr(8),1,hdr(13),hdr(9),hdr(10),hdr(11),0,0,0,0,0,0,0,0,0,0,1,hdr(12),0,LGAME_PREC_FP64,LGAME_RAD_OFF,0)

  ! grid (extents as in control.F90:101-157)
  allocate(z_scalar(ny,nx,nl),z_u(ny,nx+1,nl),z_v(ny+1,nx,nl),z_w(ny,nx,nl+1),dx(ny,nx+1,nl),dy(ny+1,nx,nl))
  allocate(dz(ny,nx,nl+1),area_x(ny,nx+1,nl),area_y(ny+1,nx,nl),area_z(ny,nx,nl+1),volume(ny,nx,nl))
  allocate(slope_x(ny,nx+1,nl),slope_y(ny+1,nx,nl),ipw(6,ny,nx,nl),area_dual_x(ny+1,nx,nl+1),area_dual_y(ny,nx+1,nl+1))
  allocate(area_dual_z(ny+1,nx+1,nl),z_area_dual_z(ny+1,nx+1,nl),fvec_x(ny+1,nx),fvec_y(ny,nx+1),fvec_z(ny+1,nx+1))
  allocate(trsk_u(3,ny),trsk_v(2,ny+1),ebg_u(ny,nx+1,nl),ebg_v(ny+1,nx,nl),ebg_w(ny,nx,nl+1))
  allocate(theta_v_bg(ny,nx,nl),exner_bg(ny,nx,nl),gravity_m_v(ny,nx,nl+1),gravity_potential(ny,nx,nl))
  read(u) z_scalar,z_u,z_v,z_w,dx,dy,dz,area_x,area_y,area_z,volume,slope_x,slope_y,ipw,area_dual_x,area_dual_y, &
          area_dual_z,z_area_dual_z,fvec_x,fvec_y,fvec_z,trsk_u,trsk_v,ebg_u,ebg_v,ebg_w,theta_v_bg,exner_bg, &
          gravity_m_v,gravity_potential
  ! states (control.F90:158-172)
  allocate(rho1(ny,nx,nl,nc),rt1(ny,nx,nl),tp1(ny,nx,nl),ep1(ny,nx,nl),u1(ny,nx+1,nl),v1(ny+1,nx,nl),w1(ny,nx,nl+1))
  allocate(rho2(ny,nx,nl,nc),rt2(ny,nx,nl),tp2(ny,nx,nl),ep2(ny,nx,nl),u2(ny,nx+1,nl),v2(ny+1,nx,nl),w2(ny,nx,nl+1))
  read(u) rho1,rt1,tp1,ep1,u1,v1,w1
  close(u)

  gv = lgame_grid_view(c_loc(z_scalar),c_loc(z_u),c_loc(z_v),c_loc(z_w),c_loc(dx),c_loc(dy),c_loc(dz), &
                       c_loc(area_x),c_loc(area_y),c_loc(area_z),c_loc(volume),c_loc(slope_x),c_loc(slope_y), &
                       c_loc(ipw),c_loc(area_dual_x),c_loc(area_dual_y),c_loc(area_dual_z),c_loc(z_area_dual_z), &
                       c_loc(fvec_x),c_loc(fvec_y),c_loc(fvec_z),c_loc(trsk_u),c_loc(trsk_v),c_loc(ebg_u), &
                       c_loc(ebg_v),c_loc(ebg_w),c_loc(theta_v_bg),c_loc(exner_bg),c_loc(gravity_m_v), &
                       c_loc(gravity_potential),c_null_ptr,c_null_ptr,c_null_ptr,c_null_ptr,c_null_ptr, &
                       c_null_ptr,c_null_ptr,c_null_ptr,c_null_ptr,c_null_ptr,c_null_ptr,c_null_ptr,0._c_double,0._c_double)
  rc = lgame_init(dims,cfg,gv,h)
  call lgame_check(rc,c_null_ptr,"lgame_init")

  ! state_2 = state_1 (control.F90:500)
  rho2 = rho1; rt2 = rt1; tp2 = tp1; ep2 = ep1; u2 = u1; v2 = v1; w2 = w1
  sv1 = lgame_state_view(c_loc(rho1),c_loc(rt1),c_loc(tp1),c_loc(ep1),c_loc(u1),c_loc(v1),c_loc(w1),c_null_ptr)
  sv2 = lgame_state_view(c_loc(rho2),c_loc(rt2),c_loc(tp2),c_loc(ep2),c_loc(u2),c_loc(v2),c_loc(w2),c_null_ptr)
  call lgame_check(lgame_set_state(h,0,sv1),h,"lgame_set_state")
  call lgame_check(lgame_set_state(h,1,sv2),h,"lgame_set_state")

  ! the ping-pong time loop of control.F90:520-524, the state staying resident on the device
  do step=0,nsteps-1
    if (mod(step,2)==0) then
      call lgame_check(lgame_pchevi(h,0,1,step,0,0._c_double),h,"lgame_pchevi")
    else
      call lgame_check(lgame_pchevi(h,1,0,step,0,0._c_double),h,"lgame_pchevi")
    endif
  enddo
  call lgame_check(lgame_sync(h),h,"lgame_sync")

  if (mod(nsteps,2)==0) then
    call lgame_check(lgame_get_state(h,0,sv1),h,"lgame_get_state")
  else
    call lgame_check(lgame_get_state(h,1,sv1),h,"lgame_get_state")
  endif
  call lgame_finalize(h)

  open(newunit=u,file=trim(fout),access="stream",form="unformatted",status="replace")
  write(u) rho1,rt1,tp1,ep1,u1,v1,w1
  close(u)
  write(*,*) "lgame_driver_demo: ",nsteps," steps done"

end program lgame_driver_demo
