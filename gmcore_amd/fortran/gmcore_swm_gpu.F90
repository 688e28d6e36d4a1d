! Shallow-water driver on the GPU path, shaped like gmcore_swm_driver
! (src/drivers/gmcore_swm_driver.F90:35-69) and gmcore_run
! (src/dynamics/gmcore_mod.F90:97-149): read &gmcore_control, build the mesh
! and filters, load the initial state, then time_integrator and the old/new
! swap of time_advance for run_steps steps.
!
! usage: gmcore_swm_gpu <namelist> <initial state file> <final state file>
! State files hold raw float64 arrays in this order: u_lon (nlon, nlat),
! v_lat (nlon, nlat-1), gz (nlon, nlat), gzs (nlon, nlat).
program gmcore_swm_gpu

  use, intrinsic :: iso_c_binding
  use gmc_kinds_mod
  use gmc_hip_mod
  use gmc_dynamics_mod
  implicit none

  type(block_type), target :: block
  character(512) :: nml_path, ic_path, out_path
  real(r8), allocatable :: u(:,:), v(:,:), gz(:,:), gzs(:,:)
  real(r8) diag(4)
  integer old, new, tmp, n, u_io

  call get_command_argument(1, nml_path)
  call get_command_argument(2, ic_path)
  call get_command_argument(3, out_path)
  call parse_namelist(trim(nml_path))
  call dynamics_init(block)

  allocate(u(num_lon, num_lat), v(num_lon, num_lat-1), gz(num_lon, num_lat), gzs(num_lon, num_lat))
  open(newunit=u_io, file=trim(ic_path), access='stream', form='unformatted', status='old')
  read(u_io) u, v, gz, gzs
  close(u_io)
  old = 1; new = 2
  call upload_field(block, old, GMC_U_LON, u)
  call upload_field(block, old, GMC_V_LAT, v)
  call upload_field(block, old, GMC_GZ, gz)
  call upload_field(block, old, GMC_GZS, gzs)
  call gmc_check(block%ctx, gmc_prepare(block%ctx, int(old, c_int32_t)), 'gmc_prepare')
  call gmc_check(block%ctx, gmc_diagnose(block%ctx, int(old, c_int32_t), diag), 'gmc_diagnose')
  write(*, '(a, es24.16, a, es24.16)') 'step 0 tm =', diag(1), ' te =', diag(2)

  do n = 1, run_steps
    call time_integrator(operators, block, old, new, dt_dyn)
    tmp = old; old = new; new = tmp
  end do

  call gmc_check(block%ctx, gmc_diagnose(block%ctx, int(old, c_int32_t), diag), 'gmc_diagnose')
  write(*, '(a, i0, a, es24.16, a, es24.16)') 'step ', run_steps, ' tm =', diag(1), ' te =', diag(2)
  call download_field(block, old, GMC_U_LON, u)
  call download_field(block, old, GMC_V_LAT, v)
  call download_field(block, old, GMC_GZ, gz)
  open(newunit=u_io, file=trim(out_path), access='stream', form='unformatted', status='replace')
  write(u_io) u, v, gz, gzs
  close(u_io)
  call dynamics_final(block)

end program gmcore_swm_gpu
