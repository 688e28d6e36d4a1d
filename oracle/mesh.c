/*
 * mesh.c -- oracle restatement of mesh_init_global and the spherical-polygon
 * areas it uses.  TEST INFRASTRUCTURE ONLY (see gmc_oracle.h).
 *
 * Follows src/dynamics/mesh_mod.F90:134-447 and
 * src/utils/sphere_geometry_mod.F90:96-122, 144-153, 302-426, with the
 * reference's mixed precision kept: sines and cosines of real(8) angles are
 * evaluated in double (cartesian_transform_1_r16, :110-122, where z is even a
 * double product), the polygon angles in quad (__float128 here, real(16)
 * there), and pi enters the quad formulas as the double constant pi
 * (const_mod.F90:12), exactly as Fortran's mixed-kind promotion does.
 */
#include <math.h>
#include <quadmath.h>
#include <string.h>
#include "gmc_oracle.h"

typedef __float128 q_t;

static const double PI = 3.141592653589793;      /* atan(1.0_r8) * 4, const_mod.F90:12 */
#define PI2 (PI * 2)
#define PI05 (PI * 0.5)
static double g_radius;                           /* const_mod radius */

/* cartesian_transform_1_r16, sphere_geometry_mod.F90:110-122 */
static void cart(double lon, double lat, q_t *x, q_t *y, q_t *z) {
  q_t cos_lat = (q_t)cos(lat);
  *x = (q_t)g_radius * cos_lat * (q_t)cos(lon);
  *y = (q_t)g_radius * cos_lat * (q_t)sin(lon);
  *z = (q_t)(g_radius * sin(lat));
}

/* inverse_cartesian_transform_1, sphere_geometry_mod.F90:136-146 */
static void icart(q_t x, q_t y, q_t z, q_t *lon, q_t *lat) {
  *lon = atan2q(y, x);
  *lat = asinq(z / (q_t)g_radius);
  if (*lon < (q_t)0.0) *lon = *lon + (q_t)PI2;
}

/* cross_product, math_mod.F90:15-25 */
static void cross(const q_t *a, const q_t *b, q_t *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

static q_t dot3(const q_t *a, const q_t *b) {
  q_t s = 0;
  s = s + a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

/* norm_vector, sphere_geometry_mod.F90:394-406 */
static void normv(q_t *v) {
  q_t s = 0;
  s = s + v[0] * v[0];
  s = s + v[1] * v[1];
  s = s + v[2] * v[2];
  q_t n = sqrtq(s);
  if (n != 0) { v[0] = v[0] / n; v[1] = v[1] / n; v[2] = v[2] / n; }
}

/* calc_sphere_angle_1, sphere_geometry_mod.F90:410-426 */
static q_t sphere_angle(const q_t *a, const q_t *b, const q_t *c) {
  q_t nab[3], nbc[3], t[3];
  cross(a, b, nab); normv(nab);
  cross(b, c, nbc); normv(nbc);
  q_t d = dot3(nab, nbc);
  if (d > (q_t)1.0) d = 1.0;
  if (d < (q_t)-1.0) d = -1.0;
  q_t res = acosq(-d);
  cross(nab, nbc, t);
  if (dot3(t, a) < 0) res = (q_t)PI2 - res;
  return res;
}

/* calc_area for a triangle, sphere_geometry_mod.F90:302-326 */
static q_t calc_area(const q_t *x, const q_t *y, const q_t *z) {
  q_t res = 0;
  for (int i = 0; i < 3; i++) {
    int im1 = i == 0 ? 2 : i - 1, ip1 = i == 2 ? 0 : i + 1;
    q_t a[3] = {x[im1], y[im1], z[im1]}, b[3] = {x[i], y[i], z[i]}, c[3] = {x[ip1], y[ip1], z[ip1]};
    res = res + sphere_angle(a, b, c);
  }
  return (q_t)(g_radius * g_radius) * (res - (q_t)(1 * PI));
}

/* calc_area_with_last_small_arc, sphere_geometry_mod.F90:328-390 */
static q_t calc_area_small_arc(const q_t *x, const q_t *y, const q_t *z) {
  q_t res = calc_area(x, y, z);
  q_t lon0, lat0, lon1, lat1, lon2, lat2;
  icart(x[0], y[0], z[0], &lon0, &lat0);
  icart(x[1], y[1], z[1], &lon1, &lat1);
  icart(x[2], y[2], z[2], &lon2, &lat2);
  if (lat1 != lat2) return -1;                 /* log_error in the reference */
  if (lat1 == 0) return res;                   /* the small arc is a great arc */
  q_t dlon = lat0 > lat1 ? lon2 - lon1 : lon1 - lon2;
  if (dlon < 0) dlon = dlon + (q_t)PI2;
  q_t xv[3], yv[3], zv[3];
  xv[0] = 0; yv[0] = 0;
  int pole_side = lat0 * lat1 >= 0 && fabsq(lat0) > fabsq(lat1);
  if (pole_side) {
    xv[1] = x[1]; yv[1] = y[1]; zv[1] = z[1];
    xv[2] = x[2]; yv[2] = y[2]; zv[2] = z[2];
  } else {
    xv[1] = x[2]; yv[1] = y[2]; zv[1] = z[2];
    xv[2] = x[1]; yv[2] = y[1]; zv[2] = z[1];
  }
  q_t area1;
  if (lat1 > 0) {
    zv[0] = g_radius;
    area1 = (q_t)(g_radius * g_radius) * dlon * ((q_t)1.0 - sinq(lat1));
  } else {
    zv[0] = -g_radius;
    area1 = (q_t)(g_radius * g_radius) * dlon * (sinq(lat1) + (q_t)1.0);
  }
  q_t area2 = calc_area(xv, yv, zv);
  q_t area3 = area1 - area2;
  /* the threshold is the default-real literal 1.0e-10 */
  if (area3 < 0 && fabsq(area3) > (q_t)(float)1.0e-10) area3 = 0;
  if (pole_side) res = res + area3; else res = res - area3;
  return res;
}

static double tri_area(int small_arc, double lon1, double lat1, double lon2, double lat2,
                       double lon3, double lat3) {
  q_t x[3], y[3], z[3];
  cart(lon1, lat1, &x[0], &y[0], &z[0]);
  cart(lon2, lat2, &x[1], &y[1], &z[1]);
  cart(lon3, lat3, &x[2], &y[2], &z[2]);
  return (double)(small_arc ? calc_area_small_arc(x, y, z) : calc_area(x, y, z));
}

int gmo_mesh_init(int nlon, int nlat, double radius, double omega, double *metrics) {
  if (nlon < 4 || nlat < 3 || !metrics) return GMC_EINVAL;
  g_radius = radius;
  const int nhalf = nlat - 1;
  memset(metrics, 0, sizeof(double) * GMC_MESH_NFIELDS * nlat);
#define MF(f, j) metrics[(f) * nlat + (j) - 1]   /* reference 1-based row j */
  const double deg = 180.0 / PI;
  const double start_lat = -PI05, end_lat = PI05;
  const double dlon = (PI2 - 0.0) / nlon;
  double full_lon[3], half_lon[3];
  for (int i = 1; i <= 2; i++) {
    full_lon[i] = 0.0 + (double)(i - 1) * dlon;
    half_lon[i] = full_lon[i] + 0.5 * dlon;
  }
  /* latitudes, mesh_mod.F90:189-237 (coarse_pole_mul = 0) */
  const double dlat0 = (end_lat - start_lat) / nhalf;
  double full_lat[nlat + 1], half_lat[nlat], full_deg[nlat + 1], half_deg[nlat];
  full_lat[1] = start_lat; full_deg[1] = start_lat * deg;
  for (int j = 2; j <= nlat - 1; j++) {
    full_lat[j] = full_lat[j - 1] + dlat0;
    if (fabs(full_lat[j]) < 1.0e-12) full_lat[j] = 0.0;
    full_deg[j] = full_lat[j] * deg;
  }
  full_lat[nlat] = end_lat; full_deg[nlat] = end_lat * deg;
  for (int j = 1; j <= nhalf; j++) {
    half_lat[j] = full_lat[j] + 0.5 * dlat0;
    if (fabs(half_lat[j]) < 1.0e-12) half_lat[j] = 0.0;
    half_deg[j] = half_lat[j] * deg;
  }
  for (int j = 1; j <= nlat; j++)
    if (full_lat[j] > 0) { full_lat[j] = -full_lat[nlat - j + 1]; full_deg[j] = -full_deg[nlat - j + 1]; }
  for (int j = 1; j <= nhalf; j++)
    if (half_lat[j] > 0) { half_lat[j] = -half_lat[nhalf - j + 1]; half_deg[j] = -half_deg[nhalf - j + 1]; }
  double fcos[nlat + 1], fsin[nlat + 1], hcos[nlat], hsin[nlat];
  for (int j = 1; j <= nhalf; j++) { hcos[j] = cos(half_lat[j]); hsin[j] = sin(half_lat[j]); }
  for (int j = 1; j <= nlat; j++) { fcos[j] = cos(full_lat[j]); fsin[j] = sin(full_lat[j]); }
  fcos[1] = 0.0; fsin[1] = -1.0; fcos[nlat] = 0.0; fsin[nlat] = 1.0;
  for (int j = 1; j <= nlat; j++) {
    MF(GMC_FULL_LAT, j) = full_lat[j]; MF(GMC_FULL_LAT_DEG, j) = full_deg[j];
    MF(GMC_FULL_COS_LAT, j) = fcos[j]; MF(GMC_FULL_SIN_LAT, j) = fsin[j];
  }
  for (int j = 1; j <= nhalf; j++) {
    MF(GMC_HALF_LAT, j) = half_lat[j]; MF(GMC_HALF_LAT_DEG, j) = half_deg[j];
    MF(GMC_HALF_COS_LAT, j) = hcos[j]; MF(GMC_HALF_SIN_LAT, j) = hsin[j];
  }
  /* cell areas and U-point sub-areas, mesh_mod.F90:269-324 */
  const double r2 = radius * radius;
  for (int j = 1; j <= nlat; j++) {
    if (j == 1) {
      MF(GMC_AREA_CELL, j) = r2 * dlon * (hsin[j] + 1.0);
      MF(GMC_AREA_SUBCELL2, j) = r2 * 0.5 * dlon * (hsin[j] + 1.0);
    } else if (j == nlat) {
      MF(GMC_AREA_CELL, j) = r2 * dlon * (1.0 - hsin[j - 1]);
      MF(GMC_AREA_SUBCELL1, j) = r2 * 0.5 * dlon * (1.0 - hsin[j - 1]);
    } else {
      MF(GMC_AREA_CELL, j) = r2 * dlon * (hsin[j] - hsin[j - 1]);
      MF(GMC_AREA_SUBCELL1, j) = r2 * 0.5 * dlon * (fsin[j] - hsin[j - 1]);
      MF(GMC_AREA_SUBCELL2, j) = r2 * 0.5 * dlon * (hsin[j] - fsin[j]);
      double w = tri_area(0, full_lon[1], full_lat[j], half_lon[1], half_lat[j - 1], half_lon[1], half_lat[j]);
      MF(GMC_AREA_LON_WEST, j) = w;
      MF(GMC_AREA_LON_EAST, j) = w;
      MF(GMC_AREA_LON, j) = w + w;
      MF(GMC_AREA_LON_NORTH, j) = tri_area(1, half_lon[1], half_lat[j], full_lon[1], full_lat[j], full_lon[2], full_lat[j]);
      MF(GMC_AREA_LON_SOUTH, j) = tri_area(1, half_lon[1], half_lat[j - 1], full_lon[2], full_lat[j], full_lon[1], full_lat[j]);
    }
  }
  /* vertex and V-point sub-areas, mesh_mod.F90:326-377 */
  for (int j = 1; j <= nhalf; j++) {
    MF(GMC_AREA_VTX, j) = r2 * dlon * (fsin[j + 1] - fsin[j]);
    double w = tri_area(0, half_lon[1], half_lat[j], full_lon[2], full_lat[j], full_lon[2], full_lat[j + 1]);
    MF(GMC_AREA_LAT_WEST, j) = w;
    MF(GMC_AREA_LAT_EAST, j) = w;
    /* the pole-side triangle of the first and last V rows is replaced by
     * the polar cap area (mesh_mod.F90:370-375), so it is not computed */
    double north = (j + 1 == nlat) ? MF(GMC_AREA_CELL, j + 1)
        : tri_area(1, full_lon[2], full_lat[j + 1], half_lon[1], half_lat[j], half_lon[2], half_lat[j]);
    double south = (j == 1) ? MF(GMC_AREA_CELL, j)
        : tri_area(1, full_lon[2], full_lat[j], half_lon[2], half_lat[j], half_lon[1], half_lat[j]);
    MF(GMC_AREA_LAT_NORTH, j) = north;
    MF(GMC_AREA_LAT_SOUTH, j) = south;
    MF(GMC_AREA_LAT, j) = north + south;
  }
  /* edge lengths, mesh_mod.F90:379-395 */
  for (int j = 2; j <= nlat - 1; j++) {
    MF(GMC_DE_LON, j) = radius * fcos[j] * dlon;
    MF(GMC_LE_LON, j) = 2.0 * MF(GMC_AREA_LON, j) / MF(GMC_DE_LON, j);
  }
  for (int j = 1; j <= nhalf; j++) {
    MF(GMC_LE_LAT, j) = radius * hcos[j] * dlon;
    MF(GMC_DE_LAT, j) = 2.0 * MF(GMC_AREA_LAT, j) / MF(GMC_LE_LAT, j);
  }
  /* 'classic' tangent weights, mesh_mod.F90:417-427; pole rows unused */
  for (int j = 2; j <= nlat - 1; j++) {
    MF(GMC_FULL_TANGENT_WGT1, j) = MF(GMC_LE_LAT, j - 1) / MF(GMC_DE_LON, j) * 0.25;
    MF(GMC_FULL_TANGENT_WGT2, j) = MF(GMC_LE_LAT, j) / MF(GMC_DE_LON, j) * 0.25;
  }
  for (int j = 1; j <= nhalf; j++) {
    MF(GMC_HALF_TANGENT_WGT1, j) = MF(GMC_LE_LON, j) / MF(GMC_DE_LAT, j) * 0.25;
    MF(GMC_HALF_TANGENT_WGT2, j) = MF(GMC_LE_LON, j + 1) / MF(GMC_DE_LAT, j) * 0.25;
  }
  /* Coriolis parameter, mesh_mod.F90:440-445 */
  for (int j = 1; j <= nlat; j++) MF(GMC_FULL_F, j) = 2.0 * omega * fsin[j];
  for (int j = 1; j <= nhalf; j++) MF(GMC_HALF_F, j) = 2.0 * omega * hsin[j];
#undef MF
  return GMC_OK;
}
