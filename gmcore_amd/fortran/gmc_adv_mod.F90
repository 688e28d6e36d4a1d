! GMCORE's tracer-advection seam with the GPU behind it.
!
! Mirrors adv_mod (src/dynamics/adv/adv_mod.F90): adv_init /
! adv_allocate_tracers build one cell-centred, non-dynamic batch
! (adv_batch_mod.F90:73-145, adv_mod.F90:334-406) whose tracers, winds and
! work arrays live on the device in libgmcore_hip.so; adv_run(batch,
! time_step) is adv_run(block, itime) (:152-312) for the winds last given
! with adv_set_wind (the u_f, v_f, mfx_lon, mfy_lat, m of block%state(itime)).
module gmc_adv_mod

  use, intrinsic :: iso_c_binding
  use gmc_kinds_mod
  use gmc_mesh_mod
  use gmc_lon_damp_mod
  use gmc_hip_mod
  implicit none

  private
  public adv_batch_type, adv_init, adv_final, adv_run, adv_set_wind, adv_set_tracer, adv_get_tracer

  type adv_batch_type
    type(c_ptr) :: adv = c_null_ptr
    integer :: nlon = 0, nlat = 0, ntracers = 0
    real(r8), allocatable :: metrics(:,:), wgt(:)
  end type adv_batch_type

contains

  subroutine adv_init(batch, ntracers)
    type(adv_batch_type), intent(inout), target :: batch
    integer, intent(in) :: ntracers
    type(gmc_mesh) mesh
    type(gmc_adv_config) cfg
    real(r8) dlon

    batch%nlon = num_lon; batch%nlat = num_lat; batch%ntracers = ntracers
    allocate(batch%metrics(num_lat, mesh_field_count), batch%wgt(num_lat))
    call mesh_metrics(num_lon, num_lat, batch%metrics, dlon)
    call lon_damp_weights(num_lat, batch%metrics, batch%wgt)
    mesh%nlon = num_lon; mesh%nlat = num_lat; mesh%dlon = dlon
    mesh%metrics = c_loc(batch%metrics)
    cfg%radius = radius
    cfg%dt_adv = dt_adv
    cfg%nstep = int(dt_adv / dt_dyn)
    cfg%ntracers = ntracers
    cfg%lon_halo_width = lon_halo_width
    select case (limiter_type)
    case ('pd');   cfg%limiter = GMC_LIMITER_PD
    case ('none'); cfg%limiter = GMC_LIMITER_NONE
    case default;  cfg%limiter = GMC_LIMITER_MONO
    end select
    cfg%lon_damp_passes = 3
    cfg%flux_type = merge(GMC_FLUX_VAN_LEER, GMC_FLUX_PPM, ffsl_flux_type == 'van_leer')
    cfg%lon_damp_wgt = c_loc(batch%wgt)
    call gmc_check(batch%adv, gmc_adv_create(mesh, cfg, batch%adv), 'gmc_adv_create')
  end subroutine adv_init

  subroutine adv_final(batch)
    type(adv_batch_type), intent(inout) :: batch
    integer(c_int) rc
    if (c_associated(batch%adv)) rc = gmc_adv_destroy(batch%adv)
    batch%adv = c_null_ptr
  end subroutine adv_final

  ! q(:,:,1,l,level) of the batch, interior (nlon, nlat)
  subroutine adv_set_tracer(batch, level, l, q)
    type(adv_batch_type), intent(inout) :: batch
    integer, intent(in) :: level, l
    real(r8), intent(in) :: q(:,:)
    call gmc_check(batch%adv, gmc_adv_upload_tracer(batch%adv, int(level, c_int32_t), int(l - 1, c_int32_t), &
                   q, int(size(q, 1), c_int64_t)), 'gmc_adv_upload_tracer')
  end subroutine adv_set_tracer

  subroutine adv_get_tracer(batch, level, l, q)
    type(adv_batch_type), intent(inout) :: batch
    integer, intent(in) :: level, l
    real(r8), intent(out) :: q(:,:)
    call gmc_check(batch%adv, gmc_adv_download_tracer(batch%adv, int(level, c_int32_t), int(l - 1, c_int32_t), &
                   q, int(size(q, 1), c_int64_t)), 'gmc_adv_download_tracer')
  end subroutine adv_get_tracer

  ! the fields of block%state(itime) adv_run reads
  subroutine adv_set_wind(batch, u_f, v_f, mfx_lon, mfy_lat, m)
    type(adv_batch_type), intent(inout) :: batch
    real(r8), intent(in), target :: u_f(:,:), v_f(:,:), mfx_lon(:,:), mfy_lat(:,:), m(:,:)
    call gmc_check(batch%adv, gmc_adv_set_wind(batch%adv, c_loc(u_f), c_loc(v_f), c_loc(mfx_lon), &
                   c_loc(mfy_lat), c_loc(m), int(size(u_f, 1), c_int64_t)), 'gmc_adv_set_wind')
  end subroutine adv_set_wind

  subroutine adv_run(batch, time_step)
    type(adv_batch_type), intent(inout) :: batch
    integer, intent(in) :: time_step
    call gmc_check(batch%adv, gmc_adv_run(batch%adv, int(time_step, c_int64_t)), 'gmc_adv_run')
  end subroutine adv_run

end module gmc_adv_mod
