/*
 * filter.c -- oracle restatement of the zonal Gaussian polar filter tables.
 * TEST INFRASTRUCTURE ONLY (see gmc_oracle.h).
 *
 * Follows src/dynamics/filter_types_mod.F90:31-47 (gaussian_weight) and
 * :49-147 (filter_init, cases 'big_filter' and 'small_filter1').
 */
#include <math.h>
#include <string.h>
#include "gmc_oracle.h"

static const double PI = 3.141592653589793;

/* gaussian_weight, filter_types_mod.F90:31-47 */
static void gaussian_weight(double width, int ngrid, double coef_b, double *w) {
  double s = coef_b * width / 2.0;
  for (int i = 1; i <= ngrid; i++) {
    int x = i - (ngrid + 1) / 2;
    w[i - 1] = exp(-((double)(x * x) / (2 * (s * s)))) / (s * sqrt(PI * 2));
  }
  double sum = 0;
  for (int i = 0; i < ngrid; i++) sum = sum + w[i];
  for (int i = 0; i < ngrid; i++) w[i] = w[i] / sum;
}

/* width and tap count of one row, filter_types_mod.F90:67-88 */
static void row_width(double dx, double dt, double lat_deg, double mws, double max_cfl,
                      double a, double c, double *width, int *n) {
  *width = 0; *n = 0;
  if (dx > 0) {
    double cfl = mws * dt / dx;
    double w = a * cfl / max_cfl * (c * (tanh(90 - fabs(lat_deg)) - 1) + 1);
    int nn = (int)ceil(w) + 2;
    if (nn % 2 == 0) nn = nn + 1;
    *width = w; *n = nn;
  }
}

/* exp_two_values (math_mod.F90:146-159) */
static double exp_two_values(double val1, double val2, double x0, double x1, double x) {
  const double w = exp((x - x0) * (x - x0) * log(1.0e-3) / ((x1 - x0) * (x1 - x0)));
  return w * val1 + (1 - w) * val2;
}

int gmo_filter_init(int nlon, int nlat, const double *metrics, double dt,
                    double max_wave_speed, double max_cfl, double coef_a,
                    double coef_b, double coef_c, double coef_d, double coef_e, int kind,
                    int32_t *ngrid_lon, int32_t *ngrid_lat, int32_t *max_taps,
                    double *wgt_lon, double *wgt_lat) {
  (void)nlon;
  if (!metrics || !ngrid_lon || !ngrid_lat || !max_taps || kind < 0 || kind > 2) return GMC_EINVAL;
  const int nhalf = nlat - 1;
  double wlon[nlat], wlat[nlat];
  memset(wlon, 0, sizeof wlon); memset(wlat, 0, sizeof wlat);
  for (int r = 0; r < nlat; r++) ngrid_lon[r] = 0;
  for (int r = 0; r < nhalf; r++) ngrid_lat[r] = 0;
  if (max_wave_speed > 0) {
    for (int r = 1; r < nlat - 1; r++) {
      int n;
      row_width(metrics[GMC_DE_LON * nlat + r], dt, metrics[GMC_FULL_LAT_DEG * nlat + r],
                max_wave_speed, max_cfl, coef_a, coef_c, &wlon[r], &n);
      ngrid_lon[r] = n;
    }
    for (int r = 0; r < nhalf; r++) {
      int n;
      row_width(metrics[GMC_LE_LAT * nlat + r], dt, metrics[GMC_HALF_LAT_DEG * nlat + r],
                max_wave_speed, max_cfl, coef_a, coef_c, &wlat[r], &n);
      ngrid_lat[r] = n;
    }
  }
  if (kind == 1) {
    /* small_filter1: taps int(d * n) made odd, width d * w (:106-124) */
    const double w = coef_d;
    for (int r = 1; r < nlat - 1; r++)
      if (ngrid_lon[r] > 1) {
        int n = (int)(w * ngrid_lon[r]);
        if (n % 2 == 0) n = n + 1;
        ngrid_lon[r] = n; wlon[r] = w * wlon[r];
      }
    for (int r = 0; r < nhalf; r++)
      if (ngrid_lat[r] > 1) {
        int n = (int)(w * ngrid_lat[r]);
        if (n % 2 == 0) n = n + 1;
        ngrid_lat[r] = n; wlat[r] = w * wlat[r];
      }
  }
  if (kind == 2) {
    /* small_filter2: the factor goes from coef_e next to the poles to about
     * coef_d at 80 degrees and beyond (filter_types_mod.F90:124-143) */
    double lat0 = -metrics[GMC_FULL_LAT_DEG * nlat + 1];
    for (int r = 1; r < nlat - 1; r++)
      if (ngrid_lon[r] > 1) {
        const double w = exp_two_values(coef_e, coef_d, lat0, 80.0, fabs(metrics[GMC_FULL_LAT_DEG * nlat + r]));
        int n = (int)(w * ngrid_lon[r]);
        if (n % 2 == 0) n = n + 1;
        ngrid_lon[r] = n; wlon[r] = w * wlon[r];
      }
    lat0 = -metrics[GMC_HALF_LAT_DEG * nlat + 0];
    for (int r = 0; r < nhalf; r++)
      if (ngrid_lat[r] > 1) {
        const double w = exp_two_values(coef_e, coef_d, lat0, 80.0, fabs(metrics[GMC_HALF_LAT_DEG * nlat + r]));
        int n = (int)(w * ngrid_lat[r]);
        if (n % 2 == 0) n = n + 1;
        ngrid_lat[r] = n; wlat[r] = w * wlat[r];
      }
  }
  int mx = 1;
  for (int r = 0; r < nlat; r++) if (ngrid_lon[r] > mx) mx = ngrid_lon[r];
  for (int r = 0; r < nhalf; r++) if (ngrid_lat[r] > mx) mx = ngrid_lat[r];
  if (!wgt_lon || !wgt_lat) { *max_taps = mx; return GMC_OK; }
  if (*max_taps < mx) return GMC_EINVAL;
  const int ld = *max_taps;
  memset(wgt_lon, 0, sizeof(double) * nlat * ld);
  memset(wgt_lat, 0, sizeof(double) * nhalf * ld);
  /* the big filter weighs rows with n > 1; the small filters every row they
   * resized (filter_types_mod.F90:93-143) */
  for (int r = 1; r < nlat - 1; r++)
    if (ngrid_lon[r] > 1 || (kind >= 1 && ngrid_lon[r] == 1 && wlon[r] > 0))
      gaussian_weight(wlon[r], ngrid_lon[r], coef_b, &wgt_lon[r * ld]);
  for (int r = 0; r < nhalf; r++)
    if (ngrid_lat[r] > 1 || (kind >= 1 && ngrid_lat[r] == 1 && wlat[r] > 0))
      gaussian_weight(wlat[r], ngrid_lat[r], coef_b, &wgt_lat[r * ld]);
  return GMC_OK;
}
