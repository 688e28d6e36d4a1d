! Host-side tables of the zonal Gaussian polar filter: per latitude row a tap
! count and normalised weights, packed as gmc_filter of include/gmcore_hip.h.
!
! Follows gaussian_weight and filter_init ('big_filter', 'small_filter1',
! 'small_filter2'), src/dynamics/filter_types_mod.F90:31-143, with
! exp_two_values (src/utils/math_mod.F90:146-159).
module gmc_filter_mod

  use gmc_kinds_mod
  use gmc_mesh_mod
  implicit none

  private
  public filter_taps, filter_weights

contains

  ! normalised Gaussian of ngrid taps and the given width
  subroutine gauss(width, ngrid, w)
    real(r8), intent(in) :: width
    integer, intent(in) :: ngrid
    real(r8), intent(out) :: w(:)
    real(r8) s
    integer k, x
    s = filter_coef_b * width / 2.0_r8
    do k = 1, ngrid
      x = k - (ngrid + 1) / 2
      w(k) = exp(-x**2 / (2 * s**2)) / (s * sqrt(pi2))
    end do
    w(1:ngrid) = w(1:ngrid) / sum(w(1:ngrid))
  end subroutine gauss

  ! CFL-based width of one row and its odd tap count (:67-88)
  subroutine row_taps(dx, lat_deg, dt, w, n)
    real(r8), intent(in) :: dx, lat_deg, dt
    real(r8), intent(out) :: w
    integer, intent(out) :: n
    real(r8) cfl
    w = 0; n = 0
    if (dx > 0) then
      cfl = max_wave_speed * dt / dx
      w = filter_coef_a * cfl / max_cfl * (filter_coef_c * (tanh(90 - abs(lat_deg)) - 1) + 1)
      n = ceiling(w) + 2
      if (mod(n, 2) == 0) n = n + 1
    end if
  end subroutine row_taps

  ! factor of small_filter2: val1 at x0, val2 * (1 - 1e-3) + val1 * 1e-3 at x1
  pure real(r8) function exp_two_values(val1, val2, x0, x1, x) result(res)
    real(r8), intent(in) :: val1, val2, x0, x1, x
    real(r8) w
    w = exp((x - x0)**2 * log(1.0e-3_r8) / (x1 - x0)**2)
    res = w * val1 + (1 - w) * val2
  end function exp_two_values

  ! Tap counts and widths.  kind 0: big_filter, 1: small_filter1,
  ! 2: small_filter2.
  subroutine filter_taps(nlat, m, dt, kind, ng_lon, ng_lat, w_lon, w_lat)
    integer, intent(in) :: nlat, kind
    real(r8), intent(in) :: m(nlat, mesh_field_count), dt
    integer, intent(out) :: ng_lon(nlat), ng_lat(nlat-1)
    real(r8), intent(out) :: w_lon(nlat), w_lat(nlat-1)
    integer j, n
    real(r8) lat0, f
    ng_lon = 0; ng_lat = 0; w_lon = 0; w_lat = 0
    if (max_wave_speed > 0) then
      do j = 2, nlat - 1
        call row_taps(m(j,mf_de_lon), m(j,mf_full_lat_deg), dt, w_lon(j), ng_lon(j))
      end do
      do j = 1, nlat - 1
        call row_taps(m(j,mf_le_lat), m(j,mf_half_lat_deg), dt, w_lat(j), ng_lat(j))
      end do
    end if
    if (kind == 1) then
      do j = 2, nlat - 1
        if (ng_lon(j) > 1) then
          n = filter_coef_d * ng_lon(j)
          if (mod(n, 2) == 0) n = n + 1
          ng_lon(j) = n; w_lon(j) = filter_coef_d * w_lon(j)
        end if
      end do
      do j = 1, nlat - 1
        if (ng_lat(j) > 1) then
          n = filter_coef_d * ng_lat(j)
          if (mod(n, 2) == 0) n = n + 1
          ng_lat(j) = n; w_lat(j) = filter_coef_d * w_lat(j)
        end if
      end do
    end if
    if (kind == 2) then
      lat0 = -m(2,mf_full_lat_deg)
      do j = 2, nlat - 1
        if (ng_lon(j) > 1) then
          f = exp_two_values(filter_coef_e, filter_coef_d, lat0, 80.0_r8, abs(m(j,mf_full_lat_deg)))
          n = f * ng_lon(j)
          if (mod(n, 2) == 0) n = n + 1
          ng_lon(j) = n; w_lon(j) = f * w_lon(j)
        end if
      end do
      lat0 = -m(1,mf_half_lat_deg)
      do j = 1, nlat - 1
        if (ng_lat(j) > 1) then
          f = exp_two_values(filter_coef_e, filter_coef_d, lat0, 80.0_r8, abs(m(j,mf_half_lat_deg)))
          n = f * ng_lat(j)
          if (mod(n, 2) == 0) n = n + 1
          ng_lat(j) = n; w_lat(j) = f * w_lat(j)
        end if
      end do
    end if
  end subroutine filter_taps

  ! Weights wl(maxn, nlat), wt(maxn, nlat-1) for the taps of filter_taps;
  ! rows with a single tap get the weight 1.
  subroutine filter_weights(nlat, maxn, ng_lon, ng_lat, w_lon, w_lat, wl, wt)
    integer, intent(in) :: nlat, maxn, ng_lon(nlat), ng_lat(nlat-1)
    real(r8), intent(in) :: w_lon(nlat), w_lat(nlat-1)
    real(r8), intent(out) :: wl(maxn, nlat), wt(maxn, nlat-1)
    integer j
    wl = 0; wt = 0
    do j = 1, nlat
      if (ng_lon(j) >= 1 .and. w_lon(j) > 0) call gauss(w_lon(j), ng_lon(j), wl(:,j))
    end do
    do j = 1, nlat - 1
      if (ng_lat(j) >= 1 .and. w_lat(j) > 0) call gauss(w_lat(j), ng_lat(j), wt(:,j))
    end do
  end subroutine filter_weights

end module gmc_filter_mod
