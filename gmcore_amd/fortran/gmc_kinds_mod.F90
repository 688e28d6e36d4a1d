! Kinds, Earth constants and the &gmcore_control keys of the dynamics path.
!
! Constants follow src/utils/const_mod.F90:12-17, 55-66.  Namelist defaults
! follow src/utils/namelist_mod.F90:10-122 and are declared the same way:
! several real(r8) keys take default-real literals (e.g. filter_coef_b = 0.4),
! so their default value is the single-precision rounding of the decimal.
module gmc_kinds_mod

  use, intrinsic :: iso_c_binding, only: c_double, c_int32_t
  implicit none

  integer, parameter :: r8 = 8
  integer, parameter :: r16 = 16

  real(r8), parameter :: pi   = atan(1.0_r8) * 4.0_r8
  real(r8), parameter :: pi2  = pi * 2
  real(r8), parameter :: pi05 = pi * 0.5_r8
  real(r8), parameter :: deg  = 180.0_r8 / pi

  real(r8) :: omega  = 2 * pi / 86400.0d0
  real(r8) :: radius = 6.37122d6
  real(r8) :: g      = 9.80616d0

  ! &gmcore_control subset
  integer  :: num_lon = 0
  integer  :: num_lat = 0
  integer  :: num_lev = 1
  real(8)  :: dt_dyn  = 0
  real(8)  :: dt_adv  = 0
  integer  :: run_steps = 0
  character(30) :: time_scheme    = 'wrfrk3'
  character(30) :: pv_scheme      = 'upwind'
  logical  :: pv_pole_stokes      = .true.
  integer  :: upwind_order_pv     = 3
  real(r8) :: upwind_wgt_pv       = 1
  integer  :: coriolis_scheme     = 1
  real(r8) :: max_wave_speed      = 300
  real(r8) :: max_cfl             = 0.5
  real(r8) :: filter_coef_a       = 1.0
  real(r8) :: filter_coef_b       = 0.4
  real(r8) :: filter_coef_c       = 0.2
  real(r8) :: filter_coef_d       = 0.375
  real(r8) :: filter_coef_e       = 0.16
  logical  :: use_div_damp        = .false.
  integer  :: div_damp_order      = 2
  real(r8) :: div_damp_coef2      = 1.0_r8 / 128.0_r8
  real(r8) :: div_damp_coef4      = 0.01_r8
  character(30) :: tvd_limiter_type = 'van_leer'
  character(256) :: initial_file  = 'N/A'
  ! tracer advection (namelist_mod.F90:38, 68-69)
  integer  :: lon_halo_width      = 2
  character(30) :: limiter_type   = 'mono'
  character(30) :: ffsl_flux_type = 'ppm'

  namelist /gmcore_control/ num_lon, num_lat, num_lev, dt_dyn, dt_adv, run_steps, &
    time_scheme, pv_scheme, pv_pole_stokes, upwind_order_pv, upwind_wgt_pv, coriolis_scheme, &
    max_wave_speed, max_cfl, filter_coef_a, filter_coef_b, filter_coef_c, filter_coef_d, filter_coef_e, &
    use_div_damp, div_damp_order, div_damp_coef2, div_damp_coef4, tvd_limiter_type, initial_file, &
    lon_halo_width, limiter_type, ffsl_flux_type

contains

  ! parse_namelist, namelist_mod.F90:219-248 (one level: SW)
  subroutine parse_namelist(file_path)
    character(*), intent(in) :: file_path
    integer u
    open(newunit=u, file=file_path, status='old')
    read(u, nml=gmcore_control)
    close(u)
    if (dt_dyn == 0) dt_dyn = dt_adv
    if (dt_adv == 0) dt_adv = dt_dyn
  end subroutine parse_namelist

end module gmc_kinds_mod
