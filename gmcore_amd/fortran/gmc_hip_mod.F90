! Thin ISO_C_BINDING layer over the C ABI of libgmcore_hip.so
! (include/gmcore_hip.h, include/gmcore_adv.h, include/gmcore_baro.h).  Every function returns 0
! on success; gmc_check
! turns a non-zero code into the message of gmc_last_error and stops, which
! keeps the reference's abort-on-error behaviour (log_error / process_stop,
! src/utils/process_mod.F90:54-68).
module gmc_hip_mod

  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: GMC_U_LON = 0, GMC_V_LAT = 1, GMC_GZ = 2, GMC_U_F = 3, &
    GMC_V_F = 4, GMC_GZ_F = 5, GMC_GZS = 6, GMC_DU = 7, GMC_DV = 8, GMC_DGZ = 9, GMC_PV = 10, &
    GMC_KE = 11, GMC_DIV = 12, GMC_M = 13
  integer(c_int32_t), parameter :: GMC_TIME_PC2 = 0, GMC_TIME_WRFRK3 = 1, GMC_TIME_EULER = 2
  integer(c_int32_t), parameter :: GMC_PV_MIDPOINT = 0, GMC_PV_UPWIND = 1, GMC_PV_TVD = 2
  integer(c_int32_t), parameter :: GMC_TVD_NONE = 0, GMC_TVD_VAN_LEER = 1, GMC_TVD_MC = 2

  type, bind(C) :: gmc_mesh
    integer(c_int32_t) :: nlon, nlat
    real(c_double) :: dlon
    type(c_ptr) :: metrics
  end type gmc_mesh

  type, bind(C) :: gmc_filter
    integer(c_int32_t) :: max_taps
    type(c_ptr) :: ngrid_lon, wgt_lon, ngrid_lat, wgt_lat
  end type gmc_filter

  type, bind(C) :: gmc_config
    real(c_double) :: g
    integer(c_int32_t) :: time_scheme, pv_scheme, upwind_order_pv
    real(c_double) :: upwind_wgt_pv
    integer(c_int32_t) :: pv_pole_stokes, coriolis_scheme, use_div_damp, div_damp_order
    real(c_double) :: div_damp_coef2
    type(gmc_filter) :: big_filter, small_filter1
    real(c_double) :: div_damp_coef4, dt_dyn
    integer(c_int32_t) :: tvd_limiter, reserved
  end type gmc_config

  integer(c_int32_t), parameter :: GMC_LIMITER_MONO = 0, GMC_LIMITER_PD = 1, GMC_LIMITER_NONE = 2
  integer(c_int32_t), parameter :: GMC_FLUX_PPM = 0, GMC_FLUX_VAN_LEER = 1

  type, bind(C) :: gmc_adv_config
    real(c_double) :: radius, dt_adv
    integer(c_int32_t) :: nstep, ntracers, lon_halo_width, limiter, lon_damp_passes, flux_type
    type(c_ptr) :: lon_damp_wgt
  end type gmc_adv_config

  ! hydrostatic column operators (include/gmcore_baro.h)
  type, bind(C) :: gmc_vcoord
    integer(c_int32_t) :: nlev, reserved
    real(c_double) :: p0, ptop
    type(c_ptr) :: hyai, hybi
  end type gmc_vcoord

  type, bind(C) :: gmc_thermo
    real(c_double) :: Rd, cpd, p0
  end type gmc_thermo

  integer(c_int32_t), parameter :: GMB_PHS = 0, GMB_GZS = 1, GMB_DPHS = 2, GMB_PT = 3, GMB_PH_LEV = 4, &
    GMB_PH_EXN_LEV = 5, GMB_PH = 6, GMB_T = 7, GMB_GZ_LEV = 8, GMB_GZ = 9, GMB_DMFDLON = 10, &
    GMB_DMFDLAT = 11, GMB_WE_LEV = 12, GMB_WE_LEV_LON = 13, GMB_WE_LEV_LAT = 14, GMB_U_F = 15, &
    GMB_V_F = 16, GMB_M_LON = 17, GMB_M_LAT = 18, GMB_WEDUDLEV = 19, GMB_WEDVDLEV = 20, &
    GMB_PGF_LON = 21, GMB_PGF_LAT = 22, GMB_CFLZ = 23, GMB_PTF_LEV = 24, GMB_DPTFDLEV = 25, &
    GMB_U_LON = 26, GMB_V_LAT = 27, GMB_ADV_U = 28, GMB_ADV_V = 29, GMB_ADV_MFX = 30, GMB_ADV_MFY = 31, &
    GMB_ADV_CFLX = 32, GMB_ADV_CFLY = 33, GMB_ADV_DIVX = 34, GMB_ADV_DIVY = 35, GMB_PTF_LON = 36, &
    GMB_PTF_LAT = 37, GMB_DPTFDLON = 38, GMB_DPTFDLAT = 39

  interface
    integer(c_int) function gmc_create(mesh, cfg, ctx) bind(C, name='gmc_create')
      import; type(gmc_mesh), intent(in) :: mesh; type(gmc_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function gmc_destroy(ctx) bind(C, name='gmc_destroy')
      import; type(c_ptr), value :: ctx
    end function
    type(c_ptr) function gmc_last_error(ctx) bind(C, name='gmc_last_error')
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function gmc_upload(ctx, level, field, host, ld) bind(C, name='gmc_upload')
      import; type(c_ptr), value :: ctx; integer(c_int32_t), value :: level, field
      real(c_double), intent(in) :: host(*); integer(c_int64_t), value :: ld
    end function
    integer(c_int) function gmc_download(ctx, level, field, host, ld) bind(C, name='gmc_download')
      import; type(c_ptr), value :: ctx; integer(c_int32_t), value :: level, field
      real(c_double), intent(out) :: host(*); integer(c_int64_t), value :: ld
    end function
    integer(c_int) function gmc_prepare(ctx, level) bind(C, name='gmc_prepare')
      import; type(c_ptr), value :: ctx; integer(c_int32_t), value :: level
    end function
    integer(c_int) function gmc_space_operators(ctx, old, star, new, tend1, tend2, dt, pass) &
        bind(C, name='gmc_space_operators')
      import; type(c_ptr), value :: ctx
      integer(c_int32_t), value :: old, star, new, tend1, tend2, pass; real(c_double), value :: dt
    end function
    integer(c_int) function gmc_update_state(ctx, tend, old, new, dt, what) bind(C, name='gmc_update_state')
      import; type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tend, old, new, what; real(c_double), value :: dt
    end function
    integer(c_int) function gmc_time_integrate(ctx, old, new, dt) bind(C, name='gmc_time_integrate')
      import; type(c_ptr), value :: ctx; integer(c_int32_t), value :: old, new; real(c_double), value :: dt
    end function
    integer(c_int) function gmc_step(ctx, dt, nsteps) bind(C, name='gmc_step')
      import; type(c_ptr), value :: ctx; real(c_double), value :: dt; integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function gmc_current_level(ctx, old) bind(C, name='gmc_current_level')
      import; type(c_ptr), value :: ctx; integer(c_int32_t), intent(out) :: old
    end function
    integer(c_int) function gmc_diagnose(ctx, level, out) bind(C, name='gmc_diagnose')
      import; type(c_ptr), value :: ctx; integer(c_int32_t), value :: level
      real(c_double), intent(out) :: out(4)
    end function
    integer(c_int) function gmc_synchronize(ctx) bind(C, name='gmc_synchronize')
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function gmc_keep_derived(ctx, on) bind(C, name='gmc_keep_derived')
      import; type(c_ptr), value :: ctx; integer(c_int32_t), value :: on
    end function
    ! h0 / r0 files (include/gmcore_io.h); paths end with c_null_char
    integer(c_int) function gmc_history_h0(ctx, path, level, time_value, time_units, dt, new_file) &
        bind(C, name='gmc_history_h0')
      import; type(c_ptr), value :: ctx; character(kind=c_char), intent(in) :: path(*), time_units(*)
      integer(c_int32_t), value :: level, new_file; real(c_double), value :: time_value, dt
    end function
    integer(c_int) function gmc_restart_write(ctx, path, level, time_value, time_units, dt_dyn, interval) &
        bind(C, name='gmc_restart_write')
      import; type(c_ptr), value :: ctx; character(kind=c_char), intent(in) :: path(*), time_units(*), interval(*)
      integer(c_int32_t), value :: level; real(c_double), value :: time_value, dt_dyn
    end function
    integer(c_int) function gmc_restart_read(ctx, path, level, time_value) bind(C, name='gmc_restart_read')
      import; type(c_ptr), value :: ctx; character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: level; real(c_double), intent(out) :: time_value
    end function
    ! FFSL tracer advection
    integer(c_int) function gmc_adv_create(mesh, cfg, adv) bind(C, name='gmc_adv_create')
      import; type(gmc_mesh), intent(in) :: mesh; type(gmc_adv_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: adv
    end function
    integer(c_int) function gmc_adv_destroy(adv) bind(C, name='gmc_adv_destroy')
      import; type(c_ptr), value :: adv
    end function
    type(c_ptr) function gmc_adv_last_error(adv) bind(C, name='gmc_adv_last_error')
      import; type(c_ptr), value :: adv
    end function
    integer(c_int) function gmc_adv_upload_tracer(adv, level, tracer, host, ld) &
        bind(C, name='gmc_adv_upload_tracer')
      import; type(c_ptr), value :: adv; integer(c_int32_t), value :: level, tracer
      real(c_double), intent(in) :: host(*); integer(c_int64_t), value :: ld
    end function
    integer(c_int) function gmc_adv_download_tracer(adv, level, tracer, host, ld) &
        bind(C, name='gmc_adv_download_tracer')
      import; type(c_ptr), value :: adv; integer(c_int32_t), value :: level, tracer
      real(c_double), intent(out) :: host(*); integer(c_int64_t), value :: ld
    end function
    integer(c_int) function gmc_adv_set_wind(adv, u, v, mfx, mfy, m, ld) bind(C, name='gmc_adv_set_wind')
      import; type(c_ptr), value :: adv, u, v, mfx, mfy, m; integer(c_int64_t), value :: ld
    end function
    integer(c_int) function gmc_adv_run(adv, time_step) bind(C, name='gmc_adv_run')
      import; type(c_ptr), value :: adv; integer(c_int64_t), value :: time_step
    end function
    integer(c_int) function gmc_adv_run_dyn(adv, ctx, level, time_step) bind(C, name='gmc_adv_run_dyn')
      import; type(c_ptr), value :: adv, ctx; integer(c_int32_t), value :: level
      integer(c_int64_t), value :: time_step
    end function
    integer(c_int) function gmc_adv_init_dyn(adv, ctx, level) bind(C, name='gmc_adv_init_dyn')
      import; type(c_ptr), value :: adv, ctx; integer(c_int32_t), value :: level
    end function
    integer(c_int) function gmc_adv_set_wind_basis(adv, u, v, ld) bind(C, name='gmc_adv_set_wind_basis')
      import; type(c_ptr), value :: adv; real(c_double), intent(in) :: u(*), v(*)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function gmc_adv_drive(adv, time_step0, scale, nsteps) bind(C, name='gmc_adv_drive')
      import; type(c_ptr), value :: adv; integer(c_int64_t), value :: time_step0
      real(c_double), intent(in) :: scale(*); integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function gmc_adv_old_level(adv, old) bind(C, name='gmc_adv_old_level')
      import; type(c_ptr), value :: adv; integer(c_int32_t), intent(out) :: old
    end function
    integer(c_int) function gmc_adv_mass(adv, level, out) bind(C, name='gmc_adv_mass')
      import; type(c_ptr), value :: adv; integer(c_int32_t), value :: level
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function gmc_adv_fix_count(adv, count) bind(C, name='gmc_adv_fix_count')
      import; type(c_ptr), value :: adv; integer(c_int64_t), intent(out) :: count
    end function
    ! hydrostatic column operators
    integer(c_int) function gmc_baro_create(mesh, vc, th, baro) bind(C, name='gmc_baro_create')
      import; type(gmc_mesh), intent(in) :: mesh; type(gmc_vcoord), intent(in) :: vc
      type(gmc_thermo), intent(in) :: th; type(c_ptr), intent(out) :: baro
    end function
    integer(c_int) function gmc_baro_destroy(baro) bind(C, name='gmc_baro_destroy')
      import; type(c_ptr), value :: baro
    end function
    type(c_ptr) function gmc_baro_last_error(baro) bind(C, name='gmc_baro_last_error')
      import; type(c_ptr), value :: baro
    end function
    integer(c_int) function gmc_baro_set(baro, field, host, ld) bind(C, name='gmc_baro_set')
      import; type(c_ptr), value :: baro; integer(c_int32_t), value :: field
      real(c_double), intent(in) :: host(*); integer(c_int64_t), value :: ld
    end function
    integer(c_int) function gmc_baro_get(baro, field, host, ld) bind(C, name='gmc_baro_get')
      import; type(c_ptr), value :: baro; integer(c_int32_t), value :: field
      real(c_double), intent(out) :: host(*); integer(c_int64_t), value :: ld
    end function
    integer(c_int) function gmc_baro_calc_ph(baro) bind(C, name='gmc_baro_calc_ph')
      import; type(c_ptr), value :: baro
    end function
    integer(c_int) function gmc_baro_calc_t(baro) bind(C, name='gmc_baro_calc_t')
      import; type(c_ptr), value :: baro
    end function
    integer(c_int) function gmc_baro_calc_gz_lev(baro) bind(C, name='gmc_baro_calc_gz_lev')
      import; type(c_ptr), value :: baro
    end function
    integer(c_int) function gmc_baro_calc_dphsdt(baro) bind(C, name='gmc_baro_calc_dphsdt')
      import; type(c_ptr), value :: baro
    end function
    integer(c_int) function gmc_baro_calc_we_lev(baro) bind(C, name='gmc_baro_calc_we_lev')
      import; type(c_ptr), value :: baro
    end function
    integer(c_int) function gmc_baro_calc_wedudlev_wedvdlev(baro) bind(C, name='gmc_baro_calc_wedudlev_wedvdlev')
      import; type(c_ptr), value :: baro
    end function
    integer(c_int) function gmc_baro_pgf_lin97(baro) bind(C, name='gmc_baro_pgf_lin97')
      import; type(c_ptr), value :: baro
    end function
    integer(c_int) function gmc_baro_calc_ptf_lev(baro, limiter) bind(C, name='gmc_baro_calc_ptf_lev')
      import; type(c_ptr), value :: baro; integer(c_int32_t), value :: limiter
    end function
    integer(c_int) function gmc_baro_calc_ptf_hflx(baro, limiter, flux_type, lon_halo_width, dt) &
        bind(C, name='gmc_baro_calc_ptf_hflx')
      import; type(c_ptr), value :: baro; integer(c_int32_t), value :: limiter, flux_type, lon_halo_width
      real(c_double), value :: dt
    end function
    integer(c_int) function gmc_baro_calc_grad_ptf(baro, limiter, flux_type, lon_halo_width, dt) &
        bind(C, name='gmc_baro_calc_grad_ptf')
      import; type(c_ptr), value :: baro; integer(c_int32_t), value :: limiter, flux_type, lon_halo_width
      real(c_double), value :: dt
    end function
    integer(c_int) function gmc_baro_set_small_filter2(baro, filter) bind(C, name='gmc_baro_set_small_filter2')
      import; type(c_ptr), value :: baro; type(gmc_filter), intent(in) :: filter
    end function
    integer(c_int) function gmc_baro_small_filter2(baro) bind(C, name='gmc_baro_small_filter2')
      import; type(c_ptr), value :: baro
    end function
    integer(c_int) function gmc_baro_synchronize(baro) bind(C, name='gmc_baro_synchronize')
      import; type(c_ptr), value :: baro
    end function
  end interface

contains

  subroutine gmc_check(ctx, rc, what)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer n
    if (rc == 0) return
    write(*, '(a, a, a, i0)') 'gmc error in ', what, ': code ', rc
    if (c_associated(ctx)) then
      if (what(1:min(7, len(what))) == 'gmc_adv') then
        call c_f_pointer(gmc_adv_last_error(ctx), msg, [512])
      else if (what(1:min(8, len(what))) == 'gmc_baro') then
        call c_f_pointer(gmc_baro_last_error(ctx), msg, [512])
      else
        call c_f_pointer(gmc_last_error(ctx), msg, [512])
      end if
      n = 0
      do while (n < 512)
        if (msg(n + 1) == c_null_char) exit
        n = n + 1
      end do
      write(*, '(512a)') msg(1:n)
    end if
    error stop 1
  end subroutine gmc_check

end module gmc_hip_mod
