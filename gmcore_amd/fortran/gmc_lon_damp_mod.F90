! Weights of the longitudinal damping of tracers near the poles.
!
! lon_damp_init (src/dynamics/damp/lon_damp_mod.F90:43-68), full-latitude
! rows: wgt(j) = exp((lat0 - |lat_j|)**2 * log(1e-3) / (lat0 - 80)**2) with
! lat0 = -full_lat_deg(2), set to zero below 1e-6.  The pole rows are never
! damped; their weight is 0 here.
module gmc_lon_damp_mod

  use gmc_kinds_mod
  use gmc_mesh_mod
  implicit none

  private
  public lon_damp_weights

contains

  subroutine lon_damp_weights(nlat, metrics, wgt)
    integer, intent(in) :: nlat
    real(r8), intent(in) :: metrics(nlat, mesh_field_count)
    real(r8), intent(out) :: wgt(nlat)
    real(r8) lat, lat0, lon_damp_lat0
    integer j

    lon_damp_lat0 = 80
    wgt = 0
    lat0 = -metrics(2, mf_full_lat_deg)
    do j = 2, nlat - 1
      lat = abs(metrics(j, mf_full_lat_deg))
      wgt(j) = exp((lat0 - lat)**2 * log(1.0e-3_r8) / (lat0 - lon_damp_lat0)**2)
      wgt(j) = merge(0.0_r8, wgt(j), wgt(j) < 1.0e-6_r8)
    end do
  end subroutine lon_damp_weights

end module gmc_lon_damp_mod
