! Host-side mesh of the dynamics path: the metric terms of mesh_type that the
! shallow-water operators read, computed once on the CPU and handed to the
! GPU library as the packed table of include/gmcore_hip.h.
!
! Follows mesh_init_global (src/dynamics/mesh_mod.F90:134-447) and the
! spherical-polygon areas of src/utils/sphere_geometry_mod.F90:302-426 with
! the same kinds: sines/cosines of real(8) angles in real(8), polygon angles
! in real(16).  Row r of the packed table is the reference's j = r + 1.
module gmc_mesh_mod

  use gmc_kinds_mod
  implicit none

  private
  public mesh_metrics, mesh_field_count
  public mf_full_lat, mf_half_lat, mf_full_lat_deg, mf_half_lat_deg
  public mf_area_cell, mf_de_lon, mf_le_lon, mf_de_lat, mf_le_lat

  ! field order of enum gmc_mesh_field (include/gmcore_hip.h), 1-based here
  integer, parameter :: mf_full_lat = 1, mf_half_lat = 2, mf_full_lat_deg = 3, mf_half_lat_deg = 4
  integer, parameter :: mf_full_cos = 5, mf_half_cos = 6, mf_full_sin = 7, mf_half_sin = 8
  integer, parameter :: mf_area_cell = 9, mf_area_lon = 10, mf_area_lon_west = 11, mf_area_lon_east = 12
  integer, parameter :: mf_area_lon_north = 13, mf_area_lon_south = 14
  integer, parameter :: mf_area_lat = 15, mf_area_lat_west = 16, mf_area_lat_east = 17
  integer, parameter :: mf_area_lat_north = 18, mf_area_lat_south = 19, mf_area_vtx = 20
  integer, parameter :: mf_area_subcell1 = 21, mf_area_subcell2 = 22
  integer, parameter :: mf_de_lon = 23, mf_le_lon = 24, mf_de_lat = 25, mf_le_lat = 26
  integer, parameter :: mf_full_f = 27, mf_half_f = 28
  integer, parameter :: mf_full_tw1 = 29, mf_full_tw2 = 30, mf_half_tw1 = 31, mf_half_tw2 = 32
  integer, parameter :: mesh_field_count = 32

contains

  ! point on the sphere from (lon, lat), as cartesian_transform_1_r16
  subroutine to_xyz(lon, lat, p)
    real(r8), intent(in) :: lon, lat
    real(r16), intent(out) :: p(3)
    real(r16) c
    c = cos(lat)
    p(1) = radius * c * cos(lon)
    p(2) = radius * c * sin(lon)
    p(3) = radius * sin(lat)
  end subroutine to_xyz

  function cross3(a, b) result(c)
    real(r16), intent(in) :: a(3), b(3)
    real(r16) c(3)
    c(1) = a(2) * b(3) - a(3) * b(2)
    c(2) = a(3) * b(1) - a(1) * b(3)
    c(3) = a(1) * b(2) - a(2) * b(1)
  end function cross3

  function unit3(a) result(c)
    real(r16), intent(in) :: a(3)
    real(r16) c(3), n
    n = sqrt(sum(a * a))
    c = a
    if (n /= 0) c = a / n
  end function unit3

  ! interior angle at b of the spherical triangle a-b-c
  real(r16) function corner_angle(a, b, c) result(res)
    real(r16), intent(in) :: a(3), b(3), c(3)
    real(r16) n1(3), n2(3), d
    n1 = unit3(cross3(a, b))
    n2 = unit3(cross3(b, c))
    d = dot_product(n1, n2)
    d = min(d, 1.0_r16)
    d = max(d, -1.0_r16)
    res = acos(-d)
    if (dot_product(cross3(n1, n2), a) < 0) res = pi2 - res
  end function corner_angle

  ! Girard area of a spherical triangle (points as rows of p)
  real(r16) function tri_area(p) result(res)
    real(r16), intent(in) :: p(3,3)
    res = 0.0_r16 + corner_angle(p(:,3), p(:,1), p(:,2))
    res = res + corner_angle(p(:,1), p(:,2), p(:,3))
    res = res + corner_angle(p(:,2), p(:,3), p(:,1))
    res = radius**2 * (res - pi)
  end function tri_area

  ! area of a triangle whose last edge (points 2-3) lies on a latitude
  ! circle: the great-circle triangle corrected by the lune between the
  ! great and the small arc (calc_area_with_last_small_arc)
  real(r16) function tri_area_small_arc(p) result(res)
    real(r16), intent(in) :: p(3,3)
    real(r16) lon(3), lat(3), dl, cap, q(3,3), lune
    logical pole_side
    integer k
    res = tri_area(p)
    do k = 1, 3
      lon(k) = atan2(p(2,k), p(1,k))
      lat(k) = asin(p(3,k) / radius)
      if (lon(k) < 0.0d0) lon(k) = lon(k) + pi2
    end do
    if (lat(2) == 0) return
    if (lat(1) > lat(2)) then
      dl = lon(3) - lon(2)
    else
      dl = lon(2) - lon(3)
    end if
    if (dl < 0) dl = dl + pi2
    pole_side = lat(1) * lat(2) >= 0 .and. abs(lat(1)) > abs(lat(2))
    q(1,1) = 0; q(2,1) = 0
    if (pole_side) then
      q(:,2) = p(:,2); q(:,3) = p(:,3)
    else
      q(:,2) = p(:,3); q(:,3) = p(:,2)
    end if
    if (lat(2) > 0) then
      q(3,1) = radius
      cap = radius**2 * dl * (1.0 - sin(lat(2)))
    else
      q(3,1) = -radius
      cap = radius**2 * dl * (sin(lat(2)) + 1.0)
    end if
    lune = cap - tri_area(q)
    if (lune < 0 .and. abs(lune) > 1.0e-10) lune = 0
    if (pole_side) then
      res = res + lune
    else
      res = res - lune
    end if
  end function tri_area_small_arc

  real(r8) function area3(small_arc, lon1, lat1, lon2, lat2, lon3, lat3) result(res)
    logical, intent(in) :: small_arc
    real(r8), intent(in) :: lon1, lat1, lon2, lat2, lon3, lat3
    real(r16) p(3,3)
    call to_xyz(lon1, lat1, p(:,1))
    call to_xyz(lon2, lat2, p(:,2))
    call to_xyz(lon3, lat3, p(:,3))
    if (small_arc) then
      res = tri_area_small_arc(p)
    else
      res = tri_area(p)
    end if
  end function area3

  ! Fill the packed metric table m(nlat, mesh_field_count) (column f holds
  ! field f).  Returns dlon.
  subroutine mesh_metrics(nlon, nlat, m, dlon)
    integer, intent(in) :: nlon, nlat
    real(r8), intent(out) :: m(nlat, mesh_field_count)
    real(r8), intent(out) :: dlon
    real(r8) flon(2), hlon(2), dlat, r2, w
    real(r8) flat(nlat), hlat(nlat-1), fdeg(nlat), hdeg(nlat-1)
    integer i, j, nh

    nh = nlat - 1
    m = 0
    dlon = (pi2 - 0.0_r8) / nlon
    do i = 1, 2
      flon(i) = 0.0_r8 + (i - 1) * dlon
      hlon(i) = flon(i) + 0.5_r8 * dlon
    end do
    ! latitudes with the poles as cell rows, then mirrored for symmetry
    dlat = (pi05 - (-pi05)) / nh
    flat(1) = -pi05; fdeg(1) = -pi05 * deg
    do j = 2, nlat - 1
      flat(j) = flat(j-1) + dlat
      if (abs(flat(j)) < 1.0e-12) flat(j) = 0.0_r8
      fdeg(j) = flat(j) * deg
    end do
    flat(nlat) = pi05; fdeg(nlat) = pi05 * deg
    do j = 1, nh
      hlat(j) = flat(j) + 0.5_r8 * dlat
      if (abs(hlat(j)) < 1.0e-12) hlat(j) = 0.0_r8
      hdeg(j) = hlat(j) * deg
    end do
    do j = 1, nlat
      if (flat(j) > 0) then
        flat(j) = -flat(nlat-j+1); fdeg(j) = -fdeg(nlat-j+1)
      end if
    end do
    do j = 1, nh
      if (hlat(j) > 0) then
        hlat(j) = -hlat(nh-j+1); hdeg(j) = -hdeg(nh-j+1)
      end if
    end do
    m(:,mf_full_lat) = flat; m(:,mf_full_lat_deg) = fdeg
    m(1:nh,mf_half_lat) = hlat; m(1:nh,mf_half_lat_deg) = hdeg
    do j = 1, nh
      m(j,mf_half_cos) = cos(hlat(j)); m(j,mf_half_sin) = sin(hlat(j))
    end do
    do j = 1, nlat
      m(j,mf_full_cos) = cos(flat(j)); m(j,mf_full_sin) = sin(flat(j))
    end do
    m(1,mf_full_cos) = 0.0_r8; m(1,mf_full_sin) = -1.0_r8
    m(nlat,mf_full_cos) = 0.0_r8; m(nlat,mf_full_sin) = 1.0_r8

    r2 = radius**2
    do j = 1, nlat
      if (j == 1) then
        m(j,mf_area_cell) = r2 * dlon * (m(j,mf_half_sin) + 1.0d0)
        m(j,mf_area_subcell2) = r2 * 0.5d0 * dlon * (m(j,mf_half_sin) + 1.0d0)
      else if (j == nlat) then
        m(j,mf_area_cell) = r2 * dlon * (1.0 - m(j-1,mf_half_sin))
        m(j,mf_area_subcell1) = r2 * 0.5d0 * dlon * (1.0d0 - m(j-1,mf_half_sin))
      else
        m(j,mf_area_cell) = r2 * dlon * (m(j,mf_half_sin) - m(j-1,mf_half_sin))
        m(j,mf_area_subcell1) = r2 * 0.5d0 * dlon * (m(j,mf_full_sin) - m(j-1,mf_half_sin))
        m(j,mf_area_subcell2) = r2 * 0.5d0 * dlon * (m(j,mf_half_sin) - m(j,mf_full_sin))
        w = area3(.false., flon(1), flat(j), hlon(1), hlat(j-1), hlon(1), hlat(j))
        m(j,mf_area_lon_west) = w
        m(j,mf_area_lon_east) = w
        m(j,mf_area_lon) = w + w
        m(j,mf_area_lon_north) = area3(.true., hlon(1), hlat(j), flon(1), flat(j), flon(2), flat(j))
        m(j,mf_area_lon_south) = area3(.true., hlon(1), hlat(j-1), flon(2), flat(j), flon(1), flat(j))
      end if
    end do
    do j = 1, nh
      m(j,mf_area_vtx) = r2 * dlon * (m(j+1,mf_full_sin) - m(j,mf_full_sin))
      w = area3(.false., hlon(1), hlat(j), flon(2), flat(j), flon(2), flat(j+1))
      m(j,mf_area_lat_west) = w
      m(j,mf_area_lat_east) = w
      if (j + 1 == nlat) then
        m(j,mf_area_lat_north) = m(j+1,mf_area_cell)
      else
        m(j,mf_area_lat_north) = area3(.true., flon(2), flat(j+1), hlon(1), hlat(j), hlon(2), hlat(j))
      end if
      if (j == 1) then
        m(j,mf_area_lat_south) = m(j,mf_area_cell)
      else
        m(j,mf_area_lat_south) = area3(.true., flon(2), flat(j), hlon(2), hlat(j), hlon(1), hlat(j))
      end if
      m(j,mf_area_lat) = m(j,mf_area_lat_north) + m(j,mf_area_lat_south)
    end do
    do j = 2, nlat - 1
      m(j,mf_de_lon) = radius * m(j,mf_full_cos) * dlon
      m(j,mf_le_lon) = 2.0d0 * m(j,mf_area_lon) / m(j,mf_de_lon)
    end do
    do j = 1, nh
      m(j,mf_le_lat) = radius * m(j,mf_half_cos) * dlon
      m(j,mf_de_lat) = 2.0d0 * m(j,mf_area_lat) / m(j,mf_le_lat)
    end do
    do j = 2, nlat - 1
      m(j,mf_full_tw1) = m(j-1,mf_le_lat) / m(j,mf_de_lon) * 0.25d0
      m(j,mf_full_tw2) = m(j,mf_le_lat) / m(j,mf_de_lon) * 0.25d0
    end do
    do j = 1, nh
      m(j,mf_half_tw1) = m(j,mf_le_lon) / m(j,mf_de_lat) * 0.25d0
      m(j,mf_half_tw2) = m(j+1,mf_le_lon) / m(j,mf_de_lat) * 0.25d0
    end do
    do j = 1, nlat
      m(j,mf_full_f) = 2.0_r8 * omega * m(j,mf_full_sin)
    end do
    do j = 1, nh
      m(j,mf_half_f) = 2.0_r8 * omega * m(j,mf_half_sin)
    end do
  end subroutine mesh_metrics

end module gmc_mesh_mod
