! Advection driver on the GPU path, shaped like gmcore_adv_driver
! (src/drivers/gmcore_adv_driver.F90:75-125): read &gmcore_control, build the
! batch, load the tracers and the wind basis, then for every step set the
! prescribed wind of t + dt_adv and call adv_run.  The prescribed flow is the
! basis times cos(pi t / T), T = 12 days, with m = 1 and the mass fluxes equal
! to the winds (deform_case1_test_set_uv, deform_test_mod.F90:129-167).
!
! usage: gmcore_adv_gpu <namelist> <input file> <output file> <ntracers>
! Input: raw float64 u_basis (nlon, nlat), v_basis (nlon, nlat-1), then the
! ntracers tracers (nlon, nlat).  Output: the tracers of the old level after
! run_steps calls of adv_run (time steps 0 .. run_steps-1).
program gmcore_adv_gpu

  use, intrinsic :: iso_c_binding
  use gmc_kinds_mod
  use gmc_hip_mod
  use gmc_adv_mod
  implicit none

  real(r8), parameter :: period = 12 * 86400
  type(adv_batch_type) :: batch
  character(512) :: nml_path, in_path, out_path, arg
  real(r8), allocatable :: ub(:,:), vb(:,:), q(:,:,:), u(:,:), v(:,:), m(:,:)
  real(r8) cos_t
  integer ntr, l, n, u_io
  integer(c_int32_t) old

  call get_command_argument(1, nml_path)
  call get_command_argument(2, in_path)
  call get_command_argument(3, out_path)
  call get_command_argument(4, arg)
  read(arg, *) ntr
  call parse_namelist(trim(nml_path))
  call adv_init(batch, ntr)

  allocate(ub(num_lon, num_lat), vb(num_lon, num_lat-1), q(num_lon, num_lat, ntr))
  allocate(u(num_lon, num_lat), v(num_lon, num_lat-1), m(num_lon, num_lat))
  open(newunit=u_io, file=trim(in_path), access='stream', form='unformatted', status='old')
  read(u_io) ub, vb, q
  close(u_io)
  do l = 1, ntr
    call adv_set_tracer(batch, 1, l, q(:,:,l))
  end do
  ub(:,1) = 0; ub(:,num_lat) = 0
  m = 1
  do n = 0, run_steps - 1
    cos_t = cos(pi * ((n + 1) * dt_adv) / period)
    u = ub * cos_t
    v = vb * cos_t
    call adv_set_wind(batch, u, v, u, v, m)
    call adv_run(batch, n)
  end do
  call gmc_check(batch%adv, gmc_adv_old_level(batch%adv, old), 'gmc_adv_old_level')
  do l = 1, ntr
    call adv_get_tracer(batch, old, l, q(:,:,l))
  end do
  open(newunit=u_io, file=trim(out_path), access='stream', form='unformatted', status='replace')
  write(u_io) q
  close(u_io)
  write(*, '(a, i0, a)') 'advected ', run_steps, ' steps'
  call adv_final(batch)

end program gmcore_adv_gpu
