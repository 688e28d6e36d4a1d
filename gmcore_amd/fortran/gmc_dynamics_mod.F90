! GMCORE's dynamics seam with the GPU behind it.
!
! Mirrors the procedure-pointer interfaces of src/dynamics/time_schemes_mod.F90
! (space_operators_interface :24-34, step_interface :36-46,
! time_integrator_interface :48-55, time_scheme_init :63-80) and the SW
! space_operators/update_state calls, with the state and tendency levels held
! on the device by libgmcore_hip.so.  A state or tendency argument is the
! level index (1..3) of block%state / block%tend.
!
! time_scheme = 'pc2', 'wrfrk3' or 'euler' points time_integrator at the
! library's fused integrator (gmc_time_integrate); the same names with the
! suffix '_ops' compose it here from space_operators and update_state, as
! step_forward_backward does (:102-118).
module gmc_dynamics_mod

  use, intrinsic :: iso_c_binding
  use gmc_kinds_mod
  use gmc_mesh_mod
  use gmc_filter_mod
  use gmc_hip_mod
  implicit none

  private
  public block_type, dynamics_init, dynamics_final, time_integrator, operators, update_state
  public all_pass, forward_pass, backward_pass, upload_field, download_field

  integer, parameter :: all_pass = 0, forward_pass = 1, backward_pass = 2

  type block_type
    type(c_ptr) :: ctx = c_null_ptr
    integer :: nlon = 0, nlat = 0
    real(r8), allocatable :: metrics(:,:)
    integer(c_int32_t), allocatable :: ng(:,:)      ! tap counts: big lon, big lat, small lon, small lat
    real(r8), allocatable :: wbig_lon(:,:), wbig_lat(:,:), wsml_lon(:,:), wsml_lat(:,:)
    integer :: time_scheme = 0
  end type block_type

  abstract interface
    subroutine space_operators_interface(block, old, star, new, tend1, tend2, dt, pass)
      import block_type
      type(block_type), intent(inout) :: block
      integer, intent(in) :: old, star, new, tend1, tend2, pass
      real(8), intent(in) :: dt
    end subroutine space_operators_interface
    subroutine time_integrator_interface(space_operators, block, old, new, dt)
      import block_type, space_operators_interface
      procedure(space_operators_interface) space_operators
      type(block_type), intent(inout) :: block
      integer, intent(in) :: old, new
      real(8), intent(in) :: dt
    end subroutine time_integrator_interface
  end interface

  procedure(time_integrator_interface), pointer :: time_integrator => null()
  procedure(space_operators_interface), pointer :: operators => null()

contains

  subroutine dynamics_init(block)
    type(block_type), intent(inout), target :: block
    type(gmc_mesh) mesh
    type(gmc_config) cfg
    real(r8) dlon, wl(num_lat), wt(num_lat-1), sl(num_lat), st(num_lat-1)
    integer maxn
    integer(c_int) rc

    block%nlon = num_lon; block%nlat = num_lat
    allocate(block%metrics(num_lat, mesh_field_count))
    call mesh_metrics(num_lon, num_lat, block%metrics, dlon)
    allocate(block%ng(num_lat, 4))
    block%ng = 0
    call filter_taps(num_lat, block%metrics, dt_dyn, 0, block%ng(:,1), block%ng(1:num_lat-1,2), wl, wt)
    call filter_taps(num_lat, block%metrics, dt_dyn, 1, block%ng(:,3), block%ng(1:num_lat-1,4), sl, st)
    maxn = max(1, maxval(block%ng))
    allocate(block%wbig_lon(maxn, num_lat), block%wbig_lat(maxn, num_lat-1))
    allocate(block%wsml_lon(maxn, num_lat), block%wsml_lat(maxn, num_lat-1))
    call filter_weights(num_lat, maxn, block%ng(:,1), block%ng(1:num_lat-1,2), wl, wt, block%wbig_lon, block%wbig_lat)
    call filter_weights(num_lat, maxn, block%ng(:,3), block%ng(1:num_lat-1,4), sl, st, block%wsml_lon, block%wsml_lat)

    mesh%nlon = num_lon; mesh%nlat = num_lat; mesh%dlon = dlon
    mesh%metrics = c_loc(block%metrics)
    cfg%g = g
    select case (time_scheme(1:max(1, index(time_scheme, '_ops') - 1)))
    case ('wrfrk3'); block%time_scheme = GMC_TIME_WRFRK3
    case ('euler');  block%time_scheme = GMC_TIME_EULER
    case default;    block%time_scheme = GMC_TIME_PC2
    end select
    cfg%time_scheme = block%time_scheme
    select case (pv_scheme)
    case ('midpoint'); cfg%pv_scheme = GMC_PV_MIDPOINT
    case ('tvd');      cfg%pv_scheme = GMC_PV_TVD
    case default;      cfg%pv_scheme = GMC_PV_UPWIND
    end select
    select case (tvd_limiter_type)
    case ('none'); cfg%tvd_limiter = GMC_TVD_NONE
    case ('mc');   cfg%tvd_limiter = GMC_TVD_MC
    case default;  cfg%tvd_limiter = GMC_TVD_VAN_LEER
    end select
    cfg%reserved = 0
    cfg%upwind_order_pv = upwind_order_pv
    cfg%upwind_wgt_pv = upwind_wgt_pv
    cfg%pv_pole_stokes = merge(1, 0, pv_pole_stokes)
    cfg%coriolis_scheme = coriolis_scheme
    cfg%use_div_damp = merge(1, 0, use_div_damp)
    cfg%div_damp_order = div_damp_order
    cfg%div_damp_coef2 = div_damp_coef2
    cfg%div_damp_coef4 = div_damp_coef4
    cfg%dt_dyn = dt_dyn
    cfg%big_filter = gmc_filter(maxn, c_loc(block%ng(1,1)), c_loc(block%wbig_lon), &
                                c_loc(block%ng(1,2)), c_loc(block%wbig_lat))
    cfg%small_filter1 = gmc_filter(maxn, c_loc(block%ng(1,3)), c_loc(block%wsml_lon), &
                                   c_loc(block%ng(1,4)), c_loc(block%wsml_lat))
    rc = gmc_create(mesh, cfg, block%ctx)
    call gmc_check(block%ctx, rc, 'gmc_create')

    operators => space_operators
    if (index(time_scheme, '_ops') > 0) then
      time_integrator => ops_integrator
    else
      time_integrator => fused_integrator
    end if
  end subroutine dynamics_init

  subroutine dynamics_final(block)
    type(block_type), intent(inout) :: block
    integer(c_int) rc
    if (c_associated(block%ctx)) rc = gmc_destroy(block%ctx)
    block%ctx = c_null_ptr
  end subroutine dynamics_final

  subroutine upload_field(block, level, field, a)
    type(block_type), intent(inout) :: block
    integer, intent(in) :: level, field
    real(r8), intent(in) :: a(:,:)   ! (nlon, rows), the interior of the reference array
    call gmc_check(block%ctx, gmc_upload(block%ctx, int(level, c_int32_t), int(field, c_int32_t), a, &
                   int(size(a, 1), c_int64_t)), 'gmc_upload')
  end subroutine upload_field

  subroutine download_field(block, level, field, a)
    type(block_type), intent(inout) :: block
    integer, intent(in) :: level, field
    real(r8), intent(out) :: a(:,:)
    call gmc_check(block%ctx, gmc_download(block%ctx, int(level, c_int32_t), int(field, c_int32_t), a, &
                   int(size(a, 1), c_int64_t)), 'gmc_download')
  end subroutine download_field

  ! space_operators (gmcore_mod.F90:343-575, SW branches) on the device
  subroutine space_operators(block, old, star, new, tend1, tend2, dt, pass)
    type(block_type), intent(inout) :: block
    integer, intent(in) :: old, star, new, tend1, tend2, pass
    real(8), intent(in) :: dt
    call gmc_check(block%ctx, gmc_space_operators(block%ctx, int(old, c_int32_t), int(star, c_int32_t), &
                   int(new, c_int32_t), int(tend1, c_int32_t), int(tend2, c_int32_t), dt, &
                   int(pass, c_int32_t)), 'gmc_space_operators')
  end subroutine space_operators

  ! update_state (time_schemes_mod.F90:120-242); after a forward pass the
  ! tendency updates gz, after a backward pass u and v
  subroutine update_state(block, tend, old, new, dt, what)
    type(block_type), intent(inout) :: block
    integer, intent(in) :: tend, old, new, what
    real(8), intent(in) :: dt
    call gmc_check(block%ctx, gmc_update_state(block%ctx, int(tend, c_int32_t), int(old, c_int32_t), &
                   int(new, c_int32_t), dt, int(what, c_int32_t)), 'gmc_update_state')
  end subroutine update_state

  ! step_forward_backward, time_schemes_mod.F90:102-118
  subroutine step_forward_backward(space_ops, block, old, star, new, tend1, tend2, dt)
    procedure(space_operators_interface) space_ops
    type(block_type), intent(inout) :: block
    integer, intent(in) :: old, star, new, tend1, tend2
    real(8), intent(in) :: dt
    call space_ops(block, old, star, new, tend1, tend2, dt, forward_pass)
    call update_state(block, tend1, old, new, dt, 1)
    call space_ops(block, old, star, new, tend2, tend1, dt, backward_pass)
    call update_state(block, tend2, old, new, dt, 2)
  end subroutine step_forward_backward

  ! pc2 / wrfrk3 / euler composed on the host (time_schemes_mod.F90:244-288)
  subroutine ops_integrator(space_ops, block, old, new, dt)
    procedure(space_operators_interface) space_ops
    type(block_type), intent(inout) :: block
    integer, intent(in) :: old, new
    real(8), intent(in) :: dt
    select case (block%time_scheme)
    case (GMC_TIME_EULER)
      call step_forward_backward(space_ops, block, old, old, new, old, new, dt)
    case (GMC_TIME_WRFRK3)
      call step_forward_backward(space_ops, block, old, old, new, old, new, dt / 3.0_r8)
      call step_forward_backward(space_ops, block, old, new, 3, old, new, dt / 2.0_r8)
      call step_forward_backward(space_ops, block, old, 3, new, old, new, dt)
    case default
      call step_forward_backward(space_ops, block, old, old, new, old, new, dt / 2.0_r8)
      call step_forward_backward(space_ops, block, old, new, 3, old, new, dt / 2.0_r8)
      call step_forward_backward(space_ops, block, old, 3, new, old, new, dt)
    end select
  end subroutine ops_integrator

  ! the whole integrator in the library's fused stage kernels
  subroutine fused_integrator(space_ops, block, old, new, dt)
    procedure(space_operators_interface) space_ops
    type(block_type), intent(inout) :: block
    integer, intent(in) :: old, new
    real(8), intent(in) :: dt
    call gmc_check(block%ctx, gmc_time_integrate(block%ctx, int(old, c_int32_t), int(new, c_int32_t), dt), &
                   'gmc_time_integrate')
  end subroutine fused_integrator

end module gmc_dynamics_mod
