! C entry points of the Fortran host setup (mesh metrics, filter tables and
! the longitudinal damping weights of the tracer advection),
! so that C, C++ and Python hosts build the same tables as the Fortran host.
module gmc_host_capi

  use, intrinsic :: iso_c_binding
  use gmc_kinds_mod
  use gmc_mesh_mod
  use gmc_filter_mod
  use gmc_lon_damp_mod
  implicit none

contains

  ! mesh_init_global metrics into the packed table metrics[f * nlat + r]
  integer(c_int) function gmch_mesh_metrics(nlon, nlat, rad, omg, metrics, dlon) &
      bind(C, name='gmch_mesh_metrics') result(rc)
    integer(c_int), value :: nlon, nlat
    real(c_double), value :: rad, omg
    real(c_double), intent(out) :: metrics(nlat, mesh_field_count)
    real(c_double), intent(out) :: dlon
    rc = 1
    if (nlon < 4 .or. nlat < 3) return
    radius = rad
    omega = omg
    call mesh_metrics(nlon, nlat, metrics, dlon)
    rc = 0
  end function gmch_mesh_metrics

  ! filter_init tables.  coefs = (max_wave_speed, max_cfl, filter_coef_a, _b,
  ! _c, _d); kind 0 = big_filter, 1 = small_filter1.  With maxn <= 0 only the
  ! tap counts are set and maxn returns the largest; then call again with
  ! wl(maxn, nlat) and wt(maxn, nlat-1).
  integer(c_int) function gmch_filter(nlat, metrics, dt, coefs, kind, ng_lon, ng_lat, maxn, wl, wt) &
      bind(C, name='gmch_filter') result(rc)
    integer(c_int), value :: nlat, kind
    real(c_double), intent(in) :: metrics(nlat, mesh_field_count)
    real(c_double), value :: dt
    real(c_double), intent(in) :: coefs(*)   ! 6 values; kind 2 reads a 7th, filter_coef_e
    integer(c_int32_t), intent(out) :: ng_lon(nlat), ng_lat(nlat-1)
    integer(c_int32_t), intent(inout) :: maxn
    real(c_double), intent(out) :: wl(*), wt(*)
    real(r8) w_lon(nlat), w_lat(nlat-1)
    integer ngl(nlat), ngt(nlat-1)
    rc = 1
    if (kind < 0 .or. kind > 2) return
    max_wave_speed = coefs(1); max_cfl = coefs(2); filter_coef_a = coefs(3)
    filter_coef_b = coefs(4); filter_coef_c = coefs(5); filter_coef_d = coefs(6)
    if (kind == 2) filter_coef_e = coefs(7)
    call filter_taps(nlat, metrics, dt, kind, ngl, ngt, w_lon, w_lat)
    ng_lon = ngl; ng_lat = ngt
    if (maxn <= 0) then
      maxn = max(1, maxval(ngl), maxval(ngt))
      rc = 0
      return
    end if
    if (maxn < max(maxval(ngl), maxval(ngt))) return
    call filter_weights(nlat, int(maxn), ngl, ngt, w_lon, w_lat, wl(1:maxn*nlat), wt(1:maxn*(nlat-1)))
    rc = 0
  end function gmch_filter

  ! lon_damp_init weights wgt(nlat) of the tracer advection
  integer(c_int) function gmch_lon_damp_weights(nlat, metrics, wgt) &
      bind(C, name='gmch_lon_damp_weights') result(rc)
    integer(c_int), value :: nlat
    real(c_double), intent(in) :: metrics(nlat, mesh_field_count)
    real(c_double), intent(out) :: wgt(nlat)
    rc = 1
    if (nlat < 5) return
    call lon_damp_weights(nlat, metrics, wgt)
    rc = 0
  end function gmch_lon_damp_weights

end module gmc_host_capi
