/*
 * adv.c -- oracle restatement of GMCORE's FFSL tracer advection (one
 * cell-centred advection batch of one block, one model level).
 * TEST INFRASTRUCTURE ONLY (see gmc_oracle.h).
 *
 * Data layout as in the reference: halo-padded arrays (1-hx:nlon+hx,
 * 1-hy:nrow+hy), hy = 2, zero-initialised, with the reference's explicit
 * fill_halo calls (periodic in longitude, cross-pole rows mirrored with a
 * half-circle shift, parallel_mod.F90:464-565 on one rank).  hx is the
 * batch's lon_halo_width plus 4, wide enough for every index the reference
 * can reach once |cflx| <= lon_halo_width holds.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gmc_oracle.h"

#define HY 2
enum { KF = 0, KH = 1 };

struct gmo_adv {
  int nlon, nlat, nhalf, hx, ncol, nrow;
  double dlon;
  gmc_adv_config cfg;
  double *mesh;            /* [GMC_MESH_NFIELDS][nlat + 2*HY], halo rows 0 */
  double *wgt;             /* wgt_full_lat, 1-based */
  /* block%state(itime) fields read by adv_run */
  double *su, *sv, *smfx, *smfy, *sm;
  /* adv_batch_type arrays (adv_batch_mod.F90:34-57) */
  double *u, *v, *mfx, *mfy, *old_m, *cflx, *cfly, *divx, *divy;
  double *qx, *qy, *mlx, *mly, *dmx, *dmy, *m6x, *m6y, *qmf_lon, *qmf_lat;
  double **q[2];           /* q(:,:,1,l,1:2) */
  int uv_step, mf_step, old, new_;
  int64_t fixes;
  int err;
};

#define MS(s, f, j) ((s)->mesh[(f) * ((s)->nlat + 2 * HY) + (j) - 1 + HY])
#define A(s, a, i, j) ((a)[(size_t)((j) - 1 + HY) * (s)->ncol + (i) - 1 + (s)->hx])

static double *falloc(const gmo_adv *s) { return calloc((size_t)s->ncol * s->nrow, sizeof(double)); }

/* fill_halo on one rank (parallel_mod.F90:464-565), as in sw.c: longitude
 * halos of rows 1..nlat, then the pole rows copied whole and the interior
 * reversed by half a circle. */
static void fill_halo(gmo_adv *s, double *a, int kind, int west, int east, int south, int north) {
  const int nlon = s->nlon, hx = s->hx, nr = kind == KF ? s->nlat : s->nhalf;
  const int mx = nlon / 2;
  if (west)
    for (int j = 1; j <= s->nlat; j++)
      for (int i = 1 - hx; i <= 0; i++) A(s, a, i, j) = A(s, a, i + nlon, j);
  if (east)
    for (int j = 1; j <= s->nlat; j++)
      for (int i = nlon + 1; i <= nlon + hx; i++) A(s, a, i, j) = A(s, a, i - nlon, j);
  double *tmp = malloc(sizeof(double) * s->ncol * HY);
#define T(i, jj) tmp[((jj) - 1) * s->ncol + (i) - 1 + hx]
  if (south) {
    const int src0 = kind == KF ? 2 : 1;
    for (int k = 0; k < HY; k++)
      for (int i = 1 - hx; i <= nlon + hx; i++) A(s, a, i, 1 - HY + k) = A(s, a, i, src0 + k);
    for (int jj = 1; jj <= HY; jj++)
      for (int i = 1 - hx; i <= nlon + hx; i++) T(i, jj) = A(s, a, i, jj - HY);
    for (int jj = 1; jj <= HY; jj++) {
      const int j = 1 - jj;
      for (int i = 1; i <= mx; i++) A(s, a, i, j) = T(i + mx, jj);
      for (int i = mx + 1; i <= nlon; i++) A(s, a, i, j) = T(i - mx, jj);
    }
  }
  if (north) {
    const int src0 = kind == KF ? nr - HY : nr - HY + 1;
    for (int k = 0; k < HY; k++)
      for (int i = 1 - hx; i <= nlon + hx; i++) A(s, a, i, nr + 1 + k) = A(s, a, i, src0 + k);
    for (int jj = 1; jj <= HY; jj++)
      for (int i = 1 - hx; i <= nlon + hx; i++) T(i, jj) = A(s, a, i, nr + jj);
    for (int jj = 1; jj <= HY; jj++) {
      const int j = nr + jj, src = HY - jj + 1;
      for (int i = 1; i <= mx; i++) A(s, a, i, j) = T(i + mx, src);
      for (int i = mx + 1; i <= nlon; i++) A(s, a, i, j) = T(i - mx, src);
    }
  }
#undef T
  free(tmp);
}
static void fill_all(gmo_adv *s, double *a, int kind) { fill_halo(s, a, kind, 1, 1, 1, 1); }

static double row_zsum(const gmo_adv *s, const double *a, int j, double sign) {
  double w[s->nlon];
  for (int i = 1; i <= s->nlon; i++) w[i - 1] = sign * A(s, a, i, j);
  return gmo_zonal_sum(w, s->nlon, 1);
}

static void fail(gmo_adv *s, int code, const char *what, int i, int j) {
  if (!s->err) {
    s->err = code;
    fprintf(stderr, "oracle adv: %s at i=%d j=%d\n", what, i, j);
  }
}

/* ---------------------------------------------------------------------- */
/* lon_damp_init (src/dynamics/damp/lon_damp_mod.F90:43-68), full rows:
 * wgt(j) = exp((lat0 - |lat_j|)^2 ln(1e-3) / (lat0 - 80)^2), lat0 = -lat_2 in
 * degrees, set to 0 below 1e-6; pole rows 0. */
int gmo_lon_damp_weights(int nlat, const double *metrics, double *wgt) {
  const double *deg = metrics + (size_t)GMC_FULL_LAT_DEG * nlat;
  const double lat0 = -deg[1], lat_damp = 80.0;
  wgt[0] = wgt[nlat - 1] = 0.0;
  for (int r = 1; r < nlat - 1; r++) {
    const double lat = fabs(deg[r]);
    double w = exp((lat0 - lat) * (lat0 - lat) * log(1.0e-3) / ((lat0 - lat_damp) * (lat0 - lat_damp)));
    wgt[r] = w < 1.0e-6 ? 0.0 : w;
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* half_lon(i) = full_lon(i) + 0.5 dlon, full_lon(i) = (i - 1) dlon
 * (mesh_mod.F90:181-183), any i of the halo-extended range */
static inline double half_lon(const gmo_adv *s, int i) { return (double)(i - 1) * s->dlon + 0.5 * s->dlon; }

/* adv_batch_accum_uv_cell (adv_batch_mod.F90:218-377) */
static void accum_uv_cell(gmo_adv *s) {
  const size_t n = (size_t)s->ncol * s->nrow;
  const int ns = s->cfg.nstep;
  if (s->uv_step == 0) {
    memcpy(s->u, s->su, n * sizeof(double));
    memcpy(s->v, s->sv, n * sizeof(double));
    s->uv_step++;  /* non-dynamic batch */
  } else if (s->uv_step == ns) {
    for (size_t k = 0; k < n; k++) {
      s->u[k] = (s->u[k] + s->su[k]) / (double)(ns + 1);
      s->v[k] = (s->v[k] + s->sv[k]) / (double)(ns + 1);
    }
    s->uv_step = 0;
  } else {
    for (size_t k = 0; k < n; k++) {
      s->u[k] = s->u[k] + s->su[k];
      s->v[k] = s->v[k] + s->sv[k];
    }
    s->uv_step++;
  }
  if (s->uv_step != 0) return;

  const double R = s->cfg.radius, dt = s->cfg.dt_adv;
  const int nlon = s->nlon, nlat = s->nlat, nhalf = s->nhalf, hx = s->hx;
  /* cflx: 4-stage Runge-Kutta departure point along the latitude circle,
   * linear interpolation of u between half-longitude points (:254-288) */
  for (int j = 2; j <= nlat - 1; j++) {
    const double dx = MS(s, GMC_DE_LON, j), cl = MS(s, GMC_FULL_COS_LAT, j);
    for (int i = 1; i <= nlon; i++) {
      const double x0 = R * half_lon(s, i) * cl;
      double uk[4], xk;
      uk[0] = A(s, s->u, i, j);
      xk = x0 + dt / 2 * uk[0];
      for (int st = 1; st < 4; st++) {
        const int l = (int)floor((double)i + (xk - x0) / dx);
        if (l < 1 - hx || l + 1 > nlon + hx) {
          fail(s, GMC_ENUMERIC, "departure point beyond the longitude halo", i, j);
          return;
        }
        uk[st] = ((R * half_lon(s, l + 1) * cl - xk) * A(s, s->u, l, j) +
                  (xk - R * half_lon(s, l) * cl) * A(s, s->u, l + 1, j)) / dx;
        xk = st < 2 ? x0 + dt / 2 * uk[st] : x0 + dt * uk[st];
      }
      const double c = (uk[0] + 2 * uk[1] + 2 * uk[2] + uk[3]) / 6 * dt / dx;
      A(s, s->cflx, i, j) = c;
      if (fabs(c) > s->cfg.lon_halo_width) fail(s, GMC_ENUMERIC, "cflx exceeds lon_halo_width", i, j);
    }
  }
  /* cfly (:290-335): the same along meridians, v frozen at the start value
   * once the departure row leaves 1..num_half_lat-1 */
  for (int j = 1; j <= nhalf; j++) {
    const double dy = MS(s, GMC_DE_LAT, j), y0 = R * MS(s, GMC_HALF_LAT, j);
    for (int i = 1; i <= nlon; i++) {
      double vk[4], yk;
      vk[0] = A(s, s->v, i, j);
      yk = y0 + dt / 2 * vk[0];
      for (int st = 1; st < 4; st++) {
        const int l = (int)floor((double)j + (yk - y0) / dy);
        if (l < 1 || l > nhalf - 1)
          vk[st] = vk[0];
        else
          vk[st] = ((R * MS(s, GMC_HALF_LAT, l + 1) - yk) * A(s, s->v, i, l) +
                    (yk - R * MS(s, GMC_HALF_LAT, l)) * A(s, s->v, i, l + 1)) / dy;
        yk = st < 2 ? y0 + dt / 2 * vk[st] : y0 + dt * vk[st];
      }
      const double c = (vk[0] + 2 * vk[1] + 2 * vk[2] + vk[3]) / 6 * dt / dy;
      A(s, s->cfly, i, j) = c;
      if (fabs(c) > 1) fail(s, GMC_ENUMERIC, "cfly exceeds 1", i, j);
    }
  }
  /* split divergences (:336-373) */
  for (int j = 2; j <= nlat - 1; j++)
    for (int i = 1; i <= nlon; i++) {
      A(s, s->divx, i, j) = (A(s, s->u, i, j) - A(s, s->u, i - 1, j)) * MS(s, GMC_LE_LON, j) /
                            MS(s, GMC_AREA_CELL, j);
      A(s, s->divy, i, j) = (A(s, s->v, i, j) * MS(s, GMC_LE_LAT, j) -
                             A(s, s->v, i, j - 1) * MS(s, GMC_LE_LAT, j - 1)) / MS(s, GMC_AREA_CELL, j);
    }
  double pole = row_zsum(s, s->v, 1, 1.0);
  pole = pole * MS(s, GMC_LE_LAT, 1) / nlon / MS(s, GMC_AREA_CELL, 1);
  for (int i = 1; i <= nlon; i++) A(s, s->divy, i, 1) = pole;
  pole = row_zsum(s, s->v, nhalf, -1.0);
  pole = pole * MS(s, GMC_LE_LAT, nhalf) / nlon / MS(s, GMC_AREA_CELL, nlat);
  for (int i = 1; i <= nlon; i++) A(s, s->divy, i, nlat) = pole;
}

/* adv_batch_accum_mf_cell (adv_batch_mod.F90:379-403) */
static void accum_mf_cell(gmo_adv *s) {
  const size_t n = (size_t)s->ncol * s->nrow;
  const int ns = s->cfg.nstep;
  if (s->mf_step == 0) {
    memcpy(s->mfx, s->smfx, n * sizeof(double));
    memcpy(s->mfy, s->smfy, n * sizeof(double));
    s->mf_step++;
  } else if (s->mf_step == ns) {
    for (size_t k = 0; k < n; k++) {
      s->mfx[k] = (s->mfx[k] + s->smfx[k]) / (double)(ns + 1);
      s->mfy[k] = (s->mfy[k] + s->smfy[k]) / (double)(ns + 1);
    }
    s->mf_step = 0;
  } else {
    for (size_t k = 0; k < n; k++) {
      s->mfx[k] = s->mfx[k] + s->smfx[k];
      s->mfy[k] = s->mfy[k] + s->smfy[k];
    }
    s->mf_step++;
  }
}

/* ---------------------------------------------------------------------- */
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }
static inline double fmin2(double a, double b) { return a < b ? a : b; }
static inline double fmax2(double a, double b) { return a > b ? a : b; }

/* slope_simple / slope_mono / slope_pd (ffsl_mod.F90:1175-1205) */
static double slope(int limiter, double fm1, double f, double fp1) {
  const double df = (fp1 - fm1) * 0.5;
  if (limiter == GMC_LIMITER_NONE) return df;
  if (limiter == GMC_LIMITER_PD) return fsign(fmin2(fabs(df), 2 * f), df);
  const double df_min = 2 * (f - fmin2(fmin2(fm1, f), fp1));
  const double df_max = 2 * (fmax2(fmax2(fm1, f), fp1) - f);
  return fsign(fmin2(fmin2(fabs(df), df_min), df_max), df);
}

/* ppm (ffsl_mod.F90:1147-1173): limited edge values fl, fr of the parabola
 * in a cell; returns fl, df = fr - fl, f6 = 6 f - 3 (fl + fr) */
static void ppm(int lim, double fm2, double fm1, double f, double fp1, double fp2, double *fl_,
                double *df_, double *f6_) {
  const double dfl = slope(lim, fm2, fm1, f);
  const double df = slope(lim, fm1, f, fp1);
  const double dfr = slope(lim, f, fp1, fp2);
  double fl = 0.5 * (fm1 + f) + (dfl - df) / 6.0;
  double fr = 0.5 * (fp1 + f) + (df - dfr) / 6.0;
  fl = f - fsign(fmin2(fabs(df), fabs(fl - f)), df);
  fr = f + fsign(fmin2(fabs(df), fabs(fr - f)), df);
  *f6_ = 6 * f - 3 * (fl + fr);
  *df_ = fr - fl;
  *fl_ = fl;
}

/* hflx_ppm_cell (ffsl_mod.F90:778-881): PPM fluxes of mx along x through
 * the U points, of my along y through the V points, for winds u, v and
 * the batch's cflx, cfly. */
static void hflx_ppm_cell(gmo_adv *s, const double *u, const double *v, const double *mx,
                          const double *my, double *mfx, double *mfy) {
  const int nlon = s->nlon, nlat = s->nlat, nhalf = s->nhalf, lim = s->cfg.limiter;
  for (int j = 1; j <= nlat; j++)
    for (int i = 1; i <= nlon; i++) {
      ppm(lim, A(s, mx, i - 2, j), A(s, mx, i - 1, j), A(s, mx, i, j), A(s, mx, i + 1, j),
          A(s, mx, i + 2, j), &A(s, s->mlx, i, j), &A(s, s->dmx, i, j), &A(s, s->m6x, i, j));
      ppm(lim, A(s, my, i, j - 2), A(s, my, i, j - 1), A(s, my, i, j), A(s, my, i, j + 1),
          A(s, my, i, j + 2), &A(s, s->mly, i, j), &A(s, s->dmy, i, j), &A(s, s->m6y, i, j));
    }
  fill_halo(s, s->mlx, KF, 1, 1, 0, 0);
  fill_halo(s, s->dmx, KF, 1, 1, 0, 0);
  fill_halo(s, s->m6x, KF, 1, 1, 0, 0);
  fill_halo(s, s->mly, KF, 0, 0, 1, 1);
  fill_halo(s, s->dmy, KF, 0, 0, 1, 1);
  fill_halo(s, s->m6y, KF, 0, 0, 1, 1);
  for (int j = 2; j <= nlat - 1; j++)
    for (int i = 1; i <= nlon; i++) {
      const double c = A(s, s->cflx, i, j);
      const int ci = (int)c;  /* int(): truncation toward zero */
      const double cf = c - ci;
      if (c > 0) {
        const int iu = i - ci;
        const double s1 = 1 - cf, s2 = 1;
        const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
        double sum = 0;
        for (int k = i + 1 - ci; k <= i; k++) sum = sum + A(s, mx, k, j);
        A(s, mfx, i, j) = A(s, u, i, j) *
                          (sum + A(s, s->mlx, iu, j) * ds1 + 0.5 * A(s, s->dmx, iu, j) * ds2 +
                           A(s, s->m6x, iu, j) * (ds2 / 2.0 - ds3 / 3.0)) / c;
      } else if (c < 0) {
        const int iu = i - ci + 1;
        const double s1 = 0, s2 = -cf;
        const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
        double sum = 0;
        for (int k = i + 1; k <= i - ci; k++) sum = sum + A(s, mx, k, j);
        A(s, mfx, i, j) = -A(s, u, i, j) *
                          (sum + A(s, s->mlx, iu, j) * ds1 + 0.5 * A(s, s->dmx, iu, j) * ds2 +
                           A(s, s->m6x, iu, j) * (ds2 / 2.0 - ds3 / 3.0)) / c;
      } else {
        A(s, mfx, i, j) = 0;
      }
    }
  for (int j = 1; j <= nhalf; j++)
    for (int i = 1; i <= nlon; i++) {
      const double c = A(s, s->cfly, i, j);
      if (c > 0) {
        const int ju = j;
        const double s1 = 1 - c, s2 = 1;
        const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
        A(s, mfy, i, j) = A(s, v, i, j) *
                          (A(s, s->mly, i, ju) * ds1 + 0.5 * A(s, s->dmy, i, ju) * ds2 +
                           A(s, s->m6y, i, ju) * (ds2 / 2.0 - ds3 / 3.0)) / c;
      } else if (c < 0) {
        const int ju = j + 1;
        const double s1 = 0, s2 = -c;
        const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
        A(s, mfy, i, j) = -A(s, v, i, j) *
                          (A(s, s->mly, i, ju) * ds1 + 0.5 * A(s, s->dmy, i, ju) * ds2 +
                           A(s, s->m6y, i, ju) * (ds2 / 2.0 - ds3 / 3.0)) / c;
      } else {
        A(s, mfy, i, j) = 0;
      }
    }
}

/* hflx_van_leer_cell (ffsl_mod.F90:556-616): van Leer fluxes of mx along x
 * (the upwind cell's limited slope plus the whole cells crossed, over
 * cflx) and of my along y (the upwind cell's value plus half its limited
 * slope times sign(cfly) - cfly), for winds u, v. */
static void hflx_van_leer_cell(gmo_adv *s, const double *u, const double *v, const double *mx,
                               const double *my, double *mfx, double *mfy) {
  const int nlon = s->nlon, nlat = s->nlat, nhalf = s->nhalf, lim = s->cfg.limiter;
  for (int j = 2; j <= nlat - 1; j++)
    for (int i = 1; i <= nlon; i++) {
      const double c = A(s, s->cflx, i, j);
      const int ci = (int)c;  /* int(): truncation toward zero */
      const double cf = c - ci;
      if (c > 0) {
        const int iu = i - ci;
        const double dm = slope(lim, A(s, mx, iu - 1, j), A(s, mx, iu, j), A(s, mx, iu + 1, j));
        double sum = 0;
        for (int k = i + 1 - ci; k <= i; k++) sum = sum + A(s, mx, k, j);
        A(s, mfx, i, j) = A(s, u, i, j) * (cf * (A(s, mx, iu, j) + dm * 0.5 * (1 - cf)) + sum) / c;
      } else if (c < 0) {
        const int iu = i - ci + 1;
        const double dm = slope(lim, A(s, mx, iu - 1, j), A(s, mx, iu, j), A(s, mx, iu + 1, j));
        double sum = 0;
        for (int k = i + 1; k <= i - ci; k++) sum = sum + A(s, mx, k, j);
        A(s, mfx, i, j) = A(s, u, i, j) * (cf * (A(s, mx, iu, j) - dm * 0.5 * (1 + cf)) - sum) / c;
      } else {
        A(s, mfx, i, j) = 0;
      }
    }
  for (int j = 1; j <= nhalf; j++)
    for (int i = 1; i <= nlon; i++) {
      const double cf = A(s, s->cfly, i, j);
      const int ju = cf > 0 ? j : j + 1;
      const double dm = slope(lim, A(s, my, i, ju - 1), A(s, my, i, ju), A(s, my, i, ju + 1));
      A(s, mfy, i, j) = A(s, v, i, j) * (A(s, my, i, ju) + dm * 0.5 * (copysign(1.0, cf) - cf));
    }
}

/* the hflx procedure pointer of ffsl_init (ffsl_mod.F90:120-129) */
static void hflx(gmo_adv *s, const double *u, const double *v, const double *mx, const double *my,
                 double *mfx, double *mfy) {
  if (s->cfg.flux_type == GMC_FLUX_VAN_LEER) hflx_van_leer_cell(s, u, v, mx, my, mfx, mfy);
  else hflx_ppm_cell(s, u, v, mx, my, mfx, mfy);
}

/* ffsl_calc_tracer_hflx_cell (ffsl_mod.F90:318-413): inner advective
 * half-steps qx, qy from the PPM fluxes of q, then the outer fluxes of qy
 * along x and qx along y. */
static void tracer_hflx_dt(gmo_adv *s, double *q, double dt) {
  const int nlon = s->nlon, nlat = s->nlat, nhalf = s->nhalf;
  double *qmfx = s->qmf_lon, *qmfy = s->qmf_lat;
  hflx(s, s->u, s->v, q, q, qmfx, qmfy);
  fill_halo(s, qmfx, KF, 1, 1, 0, 0);
  fill_halo(s, qmfy, KH, 0, 0, 1, 1);
  for (int j = 2; j <= nlat - 1; j++)
    for (int i = 1; i <= nlon; i++) {
      const double qc = A(s, q, i, j);
      A(s, s->qx, i, j) = qc - 0.5 * ((A(s, qmfx, i, j) - A(s, qmfx, i - 1, j)) * MS(s, GMC_LE_LON, j) /
                                          MS(s, GMC_AREA_CELL, j) -
                                      A(s, s->divx, i, j) * qc) * dt;
      A(s, s->qy, i, j) = qc - 0.5 * ((A(s, qmfy, i, j) * MS(s, GMC_LE_LAT, j) -
                                       A(s, qmfy, i, j - 1) * MS(s, GMC_LE_LAT, j - 1)) /
                                          MS(s, GMC_AREA_CELL, j) -
                                      A(s, s->divy, i, j) * qc) * dt;
    }
  double pole = row_zsum(s, qmfy, 1, 1.0);
  pole = pole * MS(s, GMC_LE_LAT, 1) / nlon / MS(s, GMC_AREA_CELL, 1);
  for (int i = 1; i <= nlon; i++) {
    const double qc = A(s, q, i, 1);
    A(s, s->qx, i, 1) = qc;
    A(s, s->qy, i, 1) = qc - 0.5 * (pole - A(s, s->divy, i, 1) * qc) * dt;
  }
  pole = row_zsum(s, qmfy, nhalf, 1.0);
  pole = pole * MS(s, GMC_LE_LAT, nhalf) / nlon / MS(s, GMC_AREA_CELL, nlat);
  for (int i = 1; i <= nlon; i++) {
    const double qc = A(s, q, i, nlat);
    A(s, s->qx, i, nlat) = qc;
    A(s, s->qy, i, nlat) = qc + 0.5 * (pole - A(s, s->divy, i, nlat) * qc) * dt;
  }
  fill_halo(s, s->qx, KF, 0, 0, 1, 1);
  fill_halo(s, s->qy, KF, 1, 1, 0, 0);
  hflx(s, s->mfx, s->mfy, s->qy, s->qx, qmfx, qmfy);
}
static void tracer_hflx(gmo_adv *s, double *q) { tracer_hflx_dt(s, q, s->cfg.dt_adv); }

/* lon_damp_1d, order 2 (lon_damp_mod.F90:200-241): one Jacobi sweep of
 * f + 1/4 (f_{i-1} - 2 f + f_{i+1}) blended with weight wgt */
static void lon_damp_row(gmo_adv *s, double *f, int j, double wgt) {
  if (wgt == 0) return;
  const double c = 0.25;  /* 0.5_r8**order */
  double g[s->nlon + 1];
  for (int i = 1; i <= s->nlon; i++)
    g[i] = wgt * (A(s, f, i, j) + c * (A(s, f, i - 1, j) - 2.0 * A(s, f, i, j) + A(s, f, i + 1, j))) +
           (1 - wgt) * A(s, f, i, j);
  for (int i = 1; i <= s->nlon; i++) A(s, f, i, j) = g[i];
}

static void advance_tracer(gmo_adv *s, int l, double *pre);

/* adv_run (adv_mod.F90:152-312) */
int gmo_adv_run(gmo_adv *s, int64_t time_step) {
  if (s->err) return s->err;
  accum_uv_cell(s);
  accum_mf_cell(s);
  if (s->err) return s->err;
  if (time_step > 0 && time_step % s->cfg.nstep == 0) {
    for (int l = 0; l < s->cfg.ntracers; l++) advance_tracer(s, l, NULL);
    const int t = s->old;
    s->old = s->new_;
    s->new_ = t;
  }
  memcpy(s->old_m, s->sm, (size_t)s->ncol * s->nrow * sizeof(double));
  return s->err;
}

/* the body of adv_run's tracer loop (adv_mod.F90:175-307) for tracer l;
 * with pre != NULL, pre receives the whole new-level array (halo included)
 * as the positivity fixer finds it */
static void advance_tracer(gmo_adv *s, int l, double *pre) {
  const int nlon = s->nlon, nlat = s->nlat, nhalf = s->nhalf;
  const double dt = s->cfg.dt_adv;
  {
    {
      double *qo = s->q[s->old - 1][l], *qn = s->q[s->new_ - 1][l];
      double *qmf_lon = s->qmf_lon, *qmf_lat = s->qmf_lat, *m = s->sm;
      tracer_hflx(s, qo);
      fill_halo(s, qmf_lon, KF, 1, 0, 0, 0);
      fill_halo(s, qmf_lat, KH, 0, 0, 1, 0);
      for (int j = 2; j <= nlat - 1; j++)
        for (int i = 1; i <= nlon; i++)
          A(s, qn, i, j) = A(s, s->old_m, i, j) * A(s, qo, i, j) -
                           ((A(s, qmf_lon, i, j) - A(s, qmf_lon, i - 1, j)) * MS(s, GMC_LE_LON, j) +
                            (A(s, qmf_lat, i, j) * MS(s, GMC_LE_LAT, j) -
                             A(s, qmf_lat, i, j - 1) * MS(s, GMC_LE_LAT, j - 1))) /
                               MS(s, GMC_AREA_CELL, j) * dt;
      double pole = row_zsum(s, qmf_lat, 1, 1.0);
      pole = pole * MS(s, GMC_LE_LAT, 1) / nlon / MS(s, GMC_AREA_CELL, 1) * dt;
      for (int i = 1; i <= nlon; i++) A(s, qn, i, 1) = A(s, s->old_m, i, 1) * A(s, qo, i, 1) - pole;
      pole = row_zsum(s, qmf_lat, nhalf, 1.0);
      pole = pole * MS(s, GMC_LE_LAT, nhalf) / nlon / MS(s, GMC_AREA_CELL, nlat) * dt;
      for (int i = 1; i <= nlon; i++) A(s, qn, i, nlat) = A(s, s->old_m, i, nlat) * A(s, qo, i, nlat) + pole;
      for (int j = 1; j <= nlat; j++)
        for (int i = 1; i <= nlon; i++) A(s, qn, i, j) = A(s, qn, i, j) / A(s, m, i, j);
      if (pre) memcpy(pre, qn, (size_t)s->ncol * s->nrow * sizeof(double));
      /* horizontal positivity fixer (:236-258), sequential in i; the
       * neighbours of i = 1 and i = nlon in the longitude halo hold the
       * values of the last fill_halo of this level, and edits there are
       * discarded by the fill that follows */
      for (int j = 1; j <= nlat; j++)
        for (int i = 1; i <= nlon; i++) {
          if (A(s, qn, i, j) < 0) {
            const double qm0 = A(s, qn, i, j) * A(s, m, i, j);
            const double qm1 = A(s, qn, i - 1, j) * A(s, m, i - 1, j);
            const double qm2 = A(s, qn, i + 1, j) * A(s, m, i + 1, j);
            const double qm0_half = 0.5 * qm0;
            if (qm1 >= qm0_half && qm2 >= qm0_half) {
              A(s, qn, i - 1, j) = (qm1 - qm0_half) / A(s, m, i - 1, j);
              A(s, qn, i + 1, j) = (qm2 - qm0_half) / A(s, m, i + 1, j);
            } else if (qm1 > qm0) {
              A(s, qn, i - 1, j) = (qm1 - qm0) / A(s, m, i - 1, j);
            } else if (qm2 > qm0) {
              A(s, qn, i + 1, j) = (qm2 - qm0) / A(s, m, i + 1, j);
            } else {
              fail(s, GMC_ENUMERIC, "negative tracer", i, j);
            }
            A(s, qn, i, j) = 0;
            s->fixes++;
          }
        }
      /* With one level the vertical flux loop of vflx_ppm_cell
       * (ffsl_mod.F90:883-941) is empty and qmf_lev is never written; it is
       * taken as zero, which makes :259-297 the identity on q >= 0. */
      fill_halo(s, qn, KF, 1, 1, 0, 0);
      for (int p = 0; p < s->cfg.lon_damp_passes; p++) {
        for (int j = 2; j <= nlat - 1; j++) lon_damp_row(s, qn, j, s->wgt[j]);
        fill_all(s, qn, KF);
      }
      fill_all(s, qn, KF);
    }
  }
}

/* ---------------------------------------------------------------------- */
gmo_adv *gmo_adv_create(const gmc_mesh *mesh, const gmc_adv_config *cfg) {
  if (!mesh || !cfg || mesh->nlon < 8 || mesh->nlat < 5 || (mesh->nlon & 1) || cfg->ntracers < 1 ||
      cfg->nstep < 1 || cfg->lon_halo_width < 1 || !cfg->lon_damp_wgt)
    return NULL;
  gmo_adv *s = calloc(1, sizeof(*s));
  s->nlon = mesh->nlon;
  s->nlat = mesh->nlat;
  s->nhalf = mesh->nlat - 1;
  s->dlon = mesh->dlon;
  s->cfg = *cfg;
  s->hx = cfg->lon_halo_width + 4;
  s->ncol = s->nlon + 2 * s->hx;
  s->nrow = s->nlat + 2 * HY;
  const int mr = s->nlat + 2 * HY;
  s->mesh = calloc((size_t)GMC_MESH_NFIELDS * mr, sizeof(double));
  for (int f = 0; f < GMC_MESH_NFIELDS; f++)
    for (int r = 0; r < s->nlat; r++) s->mesh[f * mr + r + HY] = mesh->metrics[f * s->nlat + r];
  s->wgt = calloc(s->nlat + 2, sizeof(double));
  for (int r = 0; r < s->nlat; r++) s->wgt[r + 1] = cfg->lon_damp_wgt[r];
  s->cfg.lon_damp_wgt = NULL;
  double **f[] = {&s->su, &s->sv, &s->smfx, &s->smfy, &s->sm, &s->u, &s->v, &s->mfx, &s->mfy,
                  &s->old_m, &s->cflx, &s->cfly, &s->divx, &s->divy, &s->qx, &s->qy, &s->mlx,
                  &s->mly, &s->dmx, &s->dmy, &s->m6x, &s->m6y, &s->qmf_lon, &s->qmf_lat};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); k++) *f[k] = falloc(s);
  for (int lv = 0; lv < 2; lv++) {
    s->q[lv] = calloc(cfg->ntracers, sizeof(double *));
    for (int l = 0; l < cfg->ntracers; l++) s->q[lv][l] = falloc(s);
  }
  s->old = 1;
  s->new_ = 2;
  return s;
}

void gmo_adv_destroy(gmo_adv *s) {
  if (!s) return;
  double *f[] = {s->su, s->sv, s->smfx, s->smfy, s->sm, s->u, s->v, s->mfx, s->mfy, s->old_m,
                 s->cflx, s->cfly, s->divx, s->divy, s->qx, s->qy, s->mlx, s->mly, s->dmx,
                 s->dmy, s->m6x, s->m6y, s->qmf_lon, s->qmf_lat};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); k++) free(f[k]);
  for (int lv = 0; lv < 2; lv++) {
    for (int l = 0; l < s->cfg.ntracers; l++) free(s->q[lv][l]);
    free(s->q[lv]);
  }
  free(s->mesh);
  free(s->wgt);
  free(s);
}

static void put(gmo_adv *s, double *a, int rows, const double *host, int64_t ld) {
  for (int r = 0; r < rows; r++)
    for (int i = 0; i < s->nlon; i++) A(s, a, i + 1, r + 1) = host[(size_t)r * ld + i];
}
static void take(const gmo_adv *s, const double *a, int rows, double *host, int64_t ld) {
  for (int r = 0; r < rows; r++)
    for (int i = 0; i < s->nlon; i++) host[(size_t)r * ld + i] = A(s, a, i + 1, r + 1);
}

int gmo_adv_set_tracer(gmo_adv *s, int level, int tracer, const double *host, int64_t ld) {
  if (level < 1 || level > 2 || tracer < 0 || tracer >= s->cfg.ntracers || ld < s->nlon) return GMC_EINVAL;
  double *q = s->q[level - 1][tracer];
  put(s, q, s->nlat, host, ld);
  fill_all(s, q, KF);
  return 0;
}

int gmo_adv_get_tracer(gmo_adv *s, int level, int tracer, double *host, int64_t ld) {
  if (level < 1 || level > 2 || tracer < 0 || tracer >= s->cfg.ntracers || ld < s->nlon) return GMC_EINVAL;
  take(s, s->q[level - 1][tracer], s->nlat, host, ld);
  return 0;
}

/* set_uv's effect on block%state(itime) (deform_test_mod.F90:129-167): u_f
 * and v_f halo-filled, mfx_lon and mfy_lat interior only, m everywhere. */
int gmo_adv_set_wind(gmo_adv *s, const double *u, const double *v, const double *mfx,
                     const double *mfy, const double *m, int64_t ld) {
  if (ld < s->nlon) return GMC_EINVAL;
  if (u) {
    put(s, s->su, s->nlat, u, ld);
    fill_all(s, s->su, KF);
  }
  if (v) {
    put(s, s->sv, s->nhalf, v, ld);
    fill_all(s, s->sv, KH);
  }
  if (mfx) put(s, s->smfx, s->nlat, mfx, ld);
  if (mfy) put(s, s->smfy, s->nhalf, mfy, ld);
  if (m) {
    put(s, s->sm, s->nlat, m, ld);
    fill_all(s, s->sm, KF);
  }
  return 0;
}

int gmo_adv_old_level(const gmo_adv *s) { return s->old; }
int64_t gmo_adv_fix_count(const gmo_adv *s) { return s->fixes; }

/* the driver's diagnose (gmcore_adv_driver.F90:103-126) for one tracer:
 * sum over j then i of q * area_cell, in that order */
double gmo_adv_mass(gmo_adv *s, int level, int tracer) {
  const double *q = s->q[level - 1][tracer];
  double qm = 0;
  for (int j = 1; j <= s->nlat; j++)
    for (int i = 1; i <= s->nlon; i++) qm = qm + A(s, q, i, j) * MS(s, GMC_AREA_CELL, j);
  return qm;
}

/* ffsl_calc_tracer_hflx_cell (ffsl_mod.F90:318-413) of one field q (nlat
 * rows, halo filled here) with the batch's u, mfx, cflx, divx, divy (nlat
 * rows) and v, mfy, cfly (nlat - 1 rows) given and the time step dt (the
 * optional dt argument, :338): the outer fluxes into qmfx (nlat rows; the
 * pole rows are those the x-flux leaves, zero here) and qmfy (nlat - 1
 * rows).  The batch's arrays of s are overwritten; its tracers are not
 * touched. */
int gmo_adv_hflx_field(gmo_adv *s, const double *q, const double *u, const double *v, const double *mfx,
                       const double *mfy, const double *cflx, const double *cfly, const double *divx,
                       const double *divy, double dt, double *qmfx, double *qmfy, int64_t ld) {
  if (ld < s->nlon) return GMC_EINVAL;
  for (int j = 0; j < s->nlat; j++)
    for (int i = 0; i < s->nlon; i++) {
      const double c = cflx[(size_t)j * ld + i];
      if (!(fabs(c) <= s->cfg.lon_halo_width)) return GMC_ENUMERIC;   /* adv_batch_mod.F90:284-287 */
    }
  double *w = falloc(s);
  put(s, w, s->nlat, q, ld);
  fill_all(s, w, KF);
  put(s, s->u, s->nlat, u, ld);
  put(s, s->v, s->nhalf, v, ld);
  put(s, s->mfx, s->nlat, mfx, ld);
  put(s, s->mfy, s->nhalf, mfy, ld);
  put(s, s->cflx, s->nlat, cflx, ld);
  put(s, s->cfly, s->nhalf, cfly, ld);
  put(s, s->divx, s->nlat, divx, ld);
  put(s, s->divy, s->nlat, divy, ld);
  memset(s->qmf_lon, 0, (size_t)s->ncol * s->nrow * sizeof(double));
  memset(s->qmf_lat, 0, (size_t)s->ncol * s->nrow * sizeof(double));
  tracer_hflx_dt(s, w, dt);
  take(s, s->qmf_lon, s->nlat, qmfx, ld);
  take(s, s->qmf_lat, s->nhalf, qmfy, ld);
  free(w);
  return 0;
}

/* interior of a batch array: 0 u, 1 v, 2 mfx, 3 mfy, 4 cflx, 5 cfly, 6 divx,
 * 7 divy, 8 qx, 9 qy, 10 qmf_lon, 11 qmf_lat, 12 old_m */
int gmo_adv_get_internal(gmo_adv *s, int index, double *host, int64_t ld) {
  double *f[] = {s->u, s->v, s->mfx, s->mfy, s->cflx, s->cfly, s->divx, s->divy,
                 s->qx, s->qy, s->qmf_lon, s->qmf_lat, s->old_m};
  const int half[] = {0, 1, 0, 1, 0, 1, 0, 0, 0, 0, 0, 1, 0};
  if (index < 0 || index > 12 || ld < s->nlon) return GMC_EINVAL;
  take(s, f[index], half[index] ? s->nhalf : s->nlat, host, ld);
  return 0;
}

/* ppm() of one five-cell stencil, for the tests */
void gmo_ppm(int limiter, const double f[5], double out[3]) {
  ppm(limiter, f[0], f[1], f[2], f[3], f[4], &out[0], &out[1], &out[2]);
}

/* ====================================================================== */
/* Tangent linear and adjoint of one adv_run step with respect to the
 * tracers, the winds held as they are (they do not depend on q).  With the
 * winds fixed the step is piecewise linear and positively homogeneous in q:
 * the slope limiters, the PPM edge clipping and the positivity fixer choose
 * among linear expressions by comparisons of q.  The tangent linear is the
 * linear map of the piece the base state selects (the forward's branches,
 * ties resolved to the first candidate as listed below); the adjoint is its
 * transpose, formed by running the step's loops backwards, halo fills
 * included. */

static inline double sgn1(double x) { return x >= 0 ? 1.0 : -1.0; }

/* d slope / d (fm1, f, fp1) of the branch the base values select (slope_*,
 * ffsl_mod.F90:1175-1205).  mono: the smallest of |df|, df_min, df_max, the
 * first of them on a tie, with the minimum / maximum of (fm1, f, fp1) taken
 * at its first occurrence. */
static void slope_jac(int lim, double fm1, double f, double fp1, double g[3]) {
  const double df = (fp1 - fm1) * 0.5;
  g[0] = -0.5;
  g[1] = 0;
  g[2] = 0.5;
  if (lim == GMC_LIMITER_NONE) return;
  const double a = fabs(df), s = copysign(1.0, df);
  if (lim == GMC_LIMITER_PD) {
    if (2 * f < a) {  /* fsign(2 f, df) */
      g[0] = 0;
      g[1] = 2 * s * sgn1(2 * f);
      g[2] = 0;
    }
    return;
  }
  const double v[3] = {fm1, f, fp1};
  int lo = 0, hi = 0;
  for (int k = 1; k < 3; k++) {
    if (v[k] < v[lo]) lo = k;
    if (v[k] > v[hi]) hi = k;
  }
  const double dmn = 2 * (f - v[lo]), dmx = 2 * (v[hi] - f);
  int pick = 0;
  double cur = a;
  if (dmn < cur) { pick = 1; cur = dmn; }
  if (dmx < cur) pick = 2;
  if (pick == 0) return;  /* fsign(|df|, df) = df */
  g[0] = g[1] = g[2] = 0;
  if (pick == 1) {        /* s * 2 (f - min) */
    g[1] += 2 * s;
    g[lo] -= 2 * s;
  } else {                /* s * 2 (max - f) */
    g[hi] += 2 * s;
    g[1] -= 2 * s;
  }
}

/* J[k][t] = d (fl, df, f6)[k] / d f[t] of ppm (ffsl_mod.F90:1147-1173) at
 * the five base values f[0..4] (fm2 .. fp2).  The edge clipping keeps
 * f -+ df when |df| is not above |fl - f| (|fr - f|). */
void gmo_ppm_jac(int lim, const double f[5], double J[3][5]) {
  double gl[3], gc[3], gr[3];
  slope_jac(lim, f[0], f[1], f[2], gl);
  slope_jac(lim, f[1], f[2], f[3], gc);
  slope_jac(lim, f[2], f[3], f[4], gr);
  const double dfl = slope(lim, f[0], f[1], f[2]);
  const double df = slope(lim, f[1], f[2], f[3]);
  const double dfr = slope(lim, f[2], f[3], f[4]);
  const double fl0 = 0.5 * (f[1] + f[2]) + (dfl - df) / 6.0;
  const double fr0 = 0.5 * (f[3] + f[2]) + (df - dfr) / 6.0;
  double Jl0[5] = {0}, Jr0[5] = {0}, Jc[5] = {0}, Jl[5], Jr[5];
  for (int k = 0; k < 3; k++) {
    Jl0[k] += gl[k] / 6.0;
    Jl0[k + 1] -= gc[k] / 6.0;
    Jr0[k + 1] += gc[k] / 6.0;
    Jr0[k + 2] -= gr[k] / 6.0;
    Jc[k + 1] = gc[k];
  }
  Jl0[1] += 0.5;
  Jl0[2] += 0.5;
  Jr0[3] += 0.5;
  Jr0[2] += 0.5;
  const double a = fabs(df), s = copysign(1.0, df);
  const int cl = fabs(fl0 - f[2]) < a, cr = fabs(fr0 - f[2]) < a;
  const double tl = s * sgn1(fl0 - f[2]), tr = s * sgn1(fr0 - f[2]);
  for (int t = 0; t < 5; t++) {
    const double e = t == 2 ? 1.0 : 0.0;
    Jl[t] = cl ? e - tl * (Jl0[t] - e) : e - Jc[t];
    Jr[t] = cr ? e + tr * (Jr0[t] - e) : e + Jc[t];
    J[0][t] = Jl[t];
    J[1][t] = Jr[t] - Jl[t];
    J[2][t] = 6 * e - 3 * (Jl[t] + Jr[t]);
  }
}

/* adjoint of fill_halo: the halo's values are added to the cells they were
 * copied from, in the reverse order of the copies, and the halo cleared */
static void fill_halo_ad(gmo_adv *s, double *a, int kind, int west, int east, int south, int north) {
  const int nlon = s->nlon, hx = s->hx, nr = kind == KF ? s->nlat : s->nhalf;
  const int mx = nlon / 2;
  if (north) {
    const int src0 = kind == KF ? nr - HY : nr - HY + 1;
    /* interior columns of row nr + jj came from row src0 + HY - jj half a
     * circle away, its halo columns from row src0 + jj - 1 unshifted */
    for (int jj = 1; jj <= HY; jj++) {
      const int j = nr + jj;
      for (int i = 1 - hx; i <= nlon + hx; i++) {
        const int in = i >= 1 && i <= nlon;
        const int is = !in ? i : (i <= mx ? i + mx : i - mx), src = in ? src0 + HY - jj : src0 + jj - 1;
        A(s, a, is, src) += A(s, a, i, j);
        A(s, a, i, j) = 0;
      }
    }
  }
  if (south) {
    /* row 1 - jj: interior from row src0 + jj - 1 shifted, halo columns
     * from row src0 + HY - jj */
    const int src0 = kind == KF ? 2 : 1;
    for (int jj = 1; jj <= HY; jj++) {
      const int j = 1 - jj;
      for (int i = 1 - hx; i <= nlon + hx; i++) {
        const int in = i >= 1 && i <= nlon;
        const int is = !in ? i : (i <= mx ? i + mx : i - mx), src = in ? src0 + jj - 1 : src0 + HY - jj;
        A(s, a, is, src) += A(s, a, i, j);
        A(s, a, i, j) = 0;
      }
    }
  }
  if (east)
    for (int j = 1; j <= s->nlat; j++)
      for (int i = nlon + 1; i <= nlon + hx; i++) {
        A(s, a, i - nlon, j) += A(s, a, i, j);
        A(s, a, i, j) = 0;
      }
  if (west)
    for (int j = 1; j <= s->nlat; j++)
      for (int i = 1 - hx; i <= 0; i++) {
        A(s, a, i + nlon, j) += A(s, a, i, j);
        A(s, a, i, j) = 0;
      }
}

struct gmo_tlad {
  double *c[6];            /* PPM coefficient (tangents or adjoints): lx, dx, 6x, ly, dy, 6y */
  double *f1x, *f1y, *f2x, *f2y, *dqx, *dqy, *dq, *dn, *pre;
};

static double *tl_new(const gmo_adv *s) { return falloc(s); }
static void tl_zero(const gmo_adv *s, double *a) { memset(a, 0, (size_t)s->ncol * s->nrow * sizeof(double)); }

/* x-flux quadrature weights of U point (i, j) (hflx_ppm_cell :831-853):
 * mfx = wu * (sum over the whole cells k0 .. k0 + n - 1 + w1 ml + w2 dm +
 * w3 m6) of the upwind cell iu */
typedef struct { int iu, k0, n; double wu, w1, w2, w3; } XFlux;
static XFlux xflux_weights(const gmo_adv *s, const double *u, int i, int j) {
  XFlux x = {0, 0, 0, 0, 0, 0, 0};
  const double c = A(s, s->cflx, i, j);
  const int ci = (int)c;
  const double cf = c - ci;
  double s1, s2;
  if (c > 0) {
    x.iu = i - ci; x.k0 = i + 1 - ci; x.n = ci; s1 = 1 - cf; s2 = 1;
    x.wu = A(s, u, i, j) / c;
  } else if (c < 0) {
    x.iu = i - ci + 1; x.k0 = i + 1; x.n = -ci; s1 = 0; s2 = -cf;
    x.wu = -A(s, u, i, j) / c;
  } else {
    return x;
  }
  const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
  x.w1 = ds1;
  x.w2 = 0.5 * ds2;
  x.w3 = ds2 / 2.0 - ds3 / 3.0;
  return x;
}
/* y-flux weights of V point (i, j) (:855-877): upwind row ju */
typedef struct { int ju; double wv, w1, w2, w3; } YFlux;
static YFlux yflux_weights(const gmo_adv *s, const double *v, int i, int j) {
  YFlux y = {j, 0, 0, 0, 0};
  const double c = A(s, s->cfly, i, j);
  double s1, s2;
  if (c > 0) {
    y.ju = j; s1 = 1 - c; s2 = 1; y.wv = A(s, v, i, j) / c;
  } else if (c < 0) {
    y.ju = j + 1; s1 = 0; s2 = -c; y.wv = -A(s, v, i, j) / c;
  } else {
    return y;
  }
  const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
  y.w1 = ds1;
  y.w2 = 0.5 * ds2;
  y.w3 = ds2 / 2.0 - ds3 / 3.0;
  return y;
}

/* tangent of hflx_ppm_cell: base mx, my (halos filled), tangents tmx, tmy
 * (halos filled) -> flux tangents fx (rows 2..nlat-1), fy (V rows) */
static void hflx_tl(gmo_adv *s, struct gmo_tlad *w, const double *u, const double *v, const double *mx,
                    const double *my, const double *tmx, const double *tmy, double *fx, double *fy) {
  const int nlon = s->nlon, nlat = s->nlat, lim = s->cfg.limiter;
  for (int j = 1; j <= nlat; j++)
    for (int i = 1; i <= nlon; i++) {
      double f[5], J[3][5];
      for (int t = 0; t < 5; t++) f[t] = A(s, mx, i - 2 + t, j);
      gmo_ppm_jac(lim, f, J);
      for (int k = 0; k < 3; k++) {
        double acc = 0;
        for (int t = 0; t < 5; t++) acc += J[k][t] * A(s, tmx, i - 2 + t, j);
        A(s, w->c[k], i, j) = acc;
      }
      for (int t = 0; t < 5; t++) f[t] = A(s, my, i, j - 2 + t);
      gmo_ppm_jac(lim, f, J);
      for (int k = 0; k < 3; k++) {
        double acc = 0;
        for (int t = 0; t < 5; t++) acc += J[k][t] * A(s, tmy, i, j - 2 + t);
        A(s, w->c[3 + k], i, j) = acc;
      }
    }
  for (int k = 0; k < 3; k++) fill_halo(s, w->c[k], KF, 1, 1, 0, 0);
  for (int k = 3; k < 6; k++) fill_halo(s, w->c[k], KF, 0, 0, 1, 1);
  for (int j = 2; j <= nlat - 1; j++)
    for (int i = 1; i <= nlon; i++) {
      const XFlux x = xflux_weights(s, u, i, j);
      double sum = 0;
      for (int k = x.k0; k < x.k0 + x.n; k++) sum += A(s, tmx, k, j);
      A(s, fx, i, j) = x.wu * (sum + x.w1 * A(s, w->c[0], x.iu, j) + x.w2 * A(s, w->c[1], x.iu, j) +
                               x.w3 * A(s, w->c[2], x.iu, j));
    }
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= nlon; i++) {
      const YFlux y = yflux_weights(s, v, i, j);
      A(s, fy, i, j) = y.wv * (y.w1 * A(s, w->c[3], i, y.ju) + y.w2 * A(s, w->c[4], i, y.ju) +
                               y.w3 * A(s, w->c[5], i, y.ju));
    }
}

/* adjoint of hflx_tl: flux adjoints fx, fy -> added to amx, amy (halos
 * included; the caller folds them back) */
static void hflx_ad(gmo_adv *s, struct gmo_tlad *w, const double *u, const double *v, const double *mx,
                    const double *my, const double *fx, const double *fy, double *amx, double *amy) {
  const int nlon = s->nlon, nlat = s->nlat, lim = s->cfg.limiter;
  for (int k = 0; k < 6; k++) tl_zero(s, w->c[k]);
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= nlon; i++) {
      const YFlux y = yflux_weights(s, v, i, j);
      const double g = y.wv * A(s, fy, i, j);
      A(s, w->c[3], i, y.ju) += y.w1 * g;
      A(s, w->c[4], i, y.ju) += y.w2 * g;
      A(s, w->c[5], i, y.ju) += y.w3 * g;
    }
  for (int j = 2; j <= nlat - 1; j++)
    for (int i = 1; i <= nlon; i++) {
      const XFlux x = xflux_weights(s, u, i, j);
      const double g = x.wu * A(s, fx, i, j);
      for (int k = x.k0; k < x.k0 + x.n; k++) A(s, amx, k, j) += g;
      A(s, w->c[0], x.iu, j) += x.w1 * g;
      A(s, w->c[1], x.iu, j) += x.w2 * g;
      A(s, w->c[2], x.iu, j) += x.w3 * g;
    }
  for (int k = 3; k < 6; k++) fill_halo_ad(s, w->c[k], KF, 0, 0, 1, 1);
  for (int k = 0; k < 3; k++) fill_halo_ad(s, w->c[k], KF, 1, 1, 0, 0);
  for (int j = 1; j <= nlat; j++)
    for (int i = 1; i <= nlon; i++) {
      double f[5], J[3][5];
      for (int t = 0; t < 5; t++) f[t] = A(s, mx, i - 2 + t, j);
      gmo_ppm_jac(lim, f, J);
      for (int t = 0; t < 5; t++)
        for (int k = 0; k < 3; k++) A(s, amx, i - 2 + t, j) += J[k][t] * A(s, w->c[k], i, j);
      for (int t = 0; t < 5; t++) f[t] = A(s, my, i, j - 2 + t);
      gmo_ppm_jac(lim, f, J);
      for (int t = 0; t < 5; t++)
        for (int k = 0; k < 3; k++) A(s, amy, i, j - 2 + t) += J[k][t] * A(s, w->c[3 + k], i, j);
    }
}

/* adjoint of lon_damp_row: the row's Jacobi sweep read columns 0 .. nlon+1
 * (halo filled) and overwrote 1 .. nlon */
static void lon_damp_row_ad(gmo_adv *s, double *a, int j, double wgt) {
  if (wgt == 0) return;
  const double c = 0.25;
  double g[s->nlon + 1];
  for (int i = 1; i <= s->nlon; i++) {
    g[i] = A(s, a, i, j);
    A(s, a, i, j) = 0;
  }
  for (int i = 1; i <= s->nlon; i++) {
    A(s, a, i - 1, j) += wgt * c * g[i];
    A(s, a, i + 1, j) += wgt * c * g[i];
    A(s, a, i, j) += (wgt * (1 - 2.0 * c) + (1 - wgt)) * g[i];
  }
}

/* the positivity fixer's choice at each cell (adv_mod.F90:236-258), found by
 * replaying it on the base values b (the new level as the fixer finds it):
 * 0 none, 1 both neighbours, 2 west, 3 east */
static void fixer_codes(gmo_adv *s, const double *pre, unsigned char *code) {
  const int nlon = s->nlon;
  double *b = malloc((size_t)s->ncol * s->nrow * sizeof(double));
  memcpy(b, pre, (size_t)s->ncol * s->nrow * sizeof(double));
  const double *m = s->sm;
  for (int j = 1; j <= s->nlat; j++)
    for (int i = 1; i <= nlon; i++) {
      unsigned char k = 0;
      if (A(s, b, i, j) < 0) {
        const double qm0 = A(s, b, i, j) * A(s, m, i, j);
        const double qm1 = A(s, b, i - 1, j) * A(s, m, i - 1, j);
        const double qm2 = A(s, b, i + 1, j) * A(s, m, i + 1, j);
        const double h = 0.5 * qm0;
        if (qm1 >= h && qm2 >= h) {
          k = 1;
          A(s, b, i - 1, j) = (qm1 - h) / A(s, m, i - 1, j);
          A(s, b, i + 1, j) = (qm2 - h) / A(s, m, i + 1, j);
        } else if (qm1 > qm0) {
          k = 2;
          A(s, b, i - 1, j) = (qm1 - qm0) / A(s, m, i - 1, j);
        } else if (qm2 > qm0) {
          k = 3;
          A(s, b, i + 1, j) = (qm2 - qm0) / A(s, m, i + 1, j);
        } else {
          k = 4;
        }
        A(s, b, i, j) = 0;
      }
      code[(size_t)(j - 1) * nlon + i - 1] = k;
    }
  free(b);
}

static struct gmo_tlad *tlad_new(const gmo_adv *s) {
  struct gmo_tlad *w = calloc(1, sizeof(*w));
  for (int k = 0; k < 6; k++) w->c[k] = tl_new(s);
  double **f[] = {&w->f1x, &w->f1y, &w->f2x, &w->f2y, &w->dqx, &w->dqy, &w->dq, &w->dn, &w->pre};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); k++) *f[k] = tl_new(s);
  return w;
}
static void tlad_free(struct gmo_tlad *w) {
  for (int k = 0; k < 6; k++) free(w->c[k]);
  double *f[] = {w->f1x, w->f1y, w->f2x, w->f2y, w->dqx, w->dqy, w->dq, w->dn, w->pre};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); k++) free(f[k]);
  free(w);
}

#define LEL(j) MS(s, GMC_LE_LON, j)
#define LAT(j) MS(s, GMC_LE_LAT, j)
#define AC(j) MS(s, GMC_AREA_CELL, j)

/* tangent of tracer l's step: w->dq (interior) -> w->dn (interior); the
 * base is the step just taken by advance_tracer (q old, s->qx, s->qy,
 * w->pre) */
static void tl_tracer(gmo_adv *s, struct gmo_tlad *w, int l) {
  const int nlon = s->nlon, nlat = s->nlat, nhalf = s->nhalf;
  const double dt = s->cfg.dt_adv;
  const double *qo = s->q[s->old - 1][l], *m = s->sm;
  double *dq = w->dq, *dqx = w->dqx, *dqy = w->dqy, *dn = w->dn;
  fill_all(s, dq, KF);
  hflx_tl(s, w, s->u, s->v, qo, qo, dq, dq, w->f1x, w->f1y);
  fill_halo(s, w->f1x, KF, 1, 1, 0, 0);
  fill_halo(s, w->f1y, KH, 0, 0, 1, 1);
  for (int j = 2; j <= nlat - 1; j++)
    for (int i = 1; i <= nlon; i++) {
      const double qc = A(s, dq, i, j);
      A(s, dqx, i, j) = qc - 0.5 * ((A(s, w->f1x, i, j) - A(s, w->f1x, i - 1, j)) * LEL(j) / AC(j) -
                                    A(s, s->divx, i, j) * qc) * dt;
      A(s, dqy, i, j) = qc - 0.5 * ((A(s, w->f1y, i, j) * LAT(j) - A(s, w->f1y, i, j - 1) * LAT(j - 1)) / AC(j) -
                                    A(s, s->divy, i, j) * qc) * dt;
    }
  double ps = 0, pn = 0;
  for (int i = 1; i <= nlon; i++) {
    ps += A(s, w->f1y, i, 1);
    pn += A(s, w->f1y, i, nhalf);
  }
  ps = ps * LAT(1) / nlon / AC(1);
  pn = pn * LAT(nhalf) / nlon / AC(nlat);
  for (int i = 1; i <= nlon; i++) {
    double qc = A(s, dq, i, 1);
    A(s, dqx, i, 1) = qc;
    A(s, dqy, i, 1) = qc - 0.5 * (ps - A(s, s->divy, i, 1) * qc) * dt;
    qc = A(s, dq, i, nlat);
    A(s, dqx, i, nlat) = qc;
    A(s, dqy, i, nlat) = qc + 0.5 * (pn - A(s, s->divy, i, nlat) * qc) * dt;
  }
  fill_halo(s, dqx, KF, 0, 0, 1, 1);
  fill_halo(s, dqy, KF, 1, 1, 0, 0);
  hflx_tl(s, w, s->mfx, s->mfy, s->qy, s->qx, dqy, dqx, w->f2x, w->f2y);
  fill_halo(s, w->f2x, KF, 1, 0, 0, 0);
  fill_halo(s, w->f2y, KH, 0, 0, 1, 0);
  tl_zero(s, dn);
  for (int j = 2; j <= nlat - 1; j++)
    for (int i = 1; i <= nlon; i++)
      A(s, dn, i, j) = A(s, s->old_m, i, j) * A(s, dq, i, j) -
                       ((A(s, w->f2x, i, j) - A(s, w->f2x, i - 1, j)) * LEL(j) +
                        (A(s, w->f2y, i, j) * LAT(j) - A(s, w->f2y, i, j - 1) * LAT(j - 1))) / AC(j) * dt;
  ps = pn = 0;
  for (int i = 1; i <= nlon; i++) {
    ps += A(s, w->f2y, i, 1);
    pn += A(s, w->f2y, i, nhalf);
  }
  ps = ps * LAT(1) / nlon / AC(1) * dt;
  pn = pn * LAT(nhalf) / nlon / AC(nlat) * dt;
  for (int i = 1; i <= nlon; i++) {
    A(s, dn, i, 1) = A(s, s->old_m, i, 1) * A(s, dq, i, 1) - ps;
    A(s, dn, i, nlat) = A(s, s->old_m, i, nlat) * A(s, dq, i, nlat) + pn;
  }
  for (int j = 1; j <= nlat; j++)
    for (int i = 1; i <= nlon; i++) A(s, dn, i, j) = A(s, dn, i, j) / A(s, m, i, j);
  unsigned char *code = malloc((size_t)nlon * nlat);
  fixer_codes(s, w->pre, code);
  for (int j = 1; j <= nlat; j++)
    for (int i = 1; i <= nlon; i++) {
      const int k = code[(size_t)(j - 1) * nlon + i - 1];
      if (!k) continue;
      const double d0 = A(s, dn, i, j) * A(s, m, i, j);
      if (k == 1 || k == 2)
        A(s, dn, i - 1, j) = (A(s, dn, i - 1, j) * A(s, m, i - 1, j) - (k == 1 ? 0.5 : 1.0) * d0) / A(s, m, i - 1, j);
      if (k == 1 || k == 3)
        A(s, dn, i + 1, j) = (A(s, dn, i + 1, j) * A(s, m, i + 1, j) - (k == 1 ? 0.5 : 1.0) * d0) / A(s, m, i + 1, j);
      A(s, dn, i, j) = 0;
    }
  free(code);
  fill_halo(s, dn, KF, 1, 1, 0, 0);
  for (int p = 0; p < s->cfg.lon_damp_passes; p++) {
    for (int j = 2; j <= nlat - 1; j++) lon_damp_row(s, dn, j, s->wgt[j]);
    fill_all(s, dn, KF);
  }
  fill_all(s, dn, KF);
}

/* adjoint of tl_tracer: w->dn (interior adjoint of the new level, halo 0)
 * -> w->dq (interior adjoint of the old level) */
static void ad_tracer(gmo_adv *s, struct gmo_tlad *w, int l) {
  const int nlon = s->nlon, nlat = s->nlat, nhalf = s->nhalf;
  const double dt = s->cfg.dt_adv;
  const double *qo = s->q[s->old - 1][l], *m = s->sm;
  double *aq = w->dq, *aqx = w->dqx, *aqy = w->dqy, *an = w->dn;
  fill_halo_ad(s, an, KF, 1, 1, 1, 1);
  for (int p = s->cfg.lon_damp_passes - 1; p >= 0; p--) {
    fill_halo_ad(s, an, KF, 1, 1, 1, 1);
    for (int j = nlat - 1; j >= 2; j--) lon_damp_row_ad(s, an, j, s->wgt[j]);
  }
  fill_halo_ad(s, an, KF, 1, 1, 0, 0);
  unsigned char *code = malloc((size_t)nlon * nlat);
  fixer_codes(s, w->pre, code);
  for (int j = nlat; j >= 1; j--)
    for (int i = nlon; i >= 1; i--) {
      const int k = code[(size_t)(j - 1) * nlon + i - 1];
      if (!k) continue;
      double g = 0;
      if (k == 1 || k == 2) g -= (k == 1 ? 0.5 : 1.0) * A(s, m, i, j) / A(s, m, i - 1, j) * A(s, an, i - 1, j);
      if (k == 1 || k == 3) g -= (k == 1 ? 0.5 : 1.0) * A(s, m, i, j) / A(s, m, i + 1, j) * A(s, an, i + 1, j);
      A(s, an, i, j) = g;
    }
  free(code);
  /* the fixer's writes into the halo columns were overwritten: their
   * adjoints are zero */
  for (int j = 1; j <= nlat; j++) {
    A(s, an, 0, j) = 0;
    A(s, an, nlon + 1, j) = 0;
  }
  tl_zero(s, aq);
  tl_zero(s, w->f2x);
  tl_zero(s, w->f2y);
  for (int j = 1; j <= nlat; j++)
    for (int i = 1; i <= nlon; i++) A(s, an, i, j) = A(s, an, i, j) / A(s, m, i, j);
  double ps = 0, pn = 0;
  for (int i = 1; i <= nlon; i++) {
    A(s, aq, i, 1) += A(s, s->old_m, i, 1) * A(s, an, i, 1);
    A(s, aq, i, nlat) += A(s, s->old_m, i, nlat) * A(s, an, i, nlat);
    ps += A(s, an, i, 1);
    pn += A(s, an, i, nlat);
  }
  for (int i = 1; i <= nlon; i++) {
    A(s, w->f2y, i, 1) -= LAT(1) / nlon / AC(1) * dt * ps;
    A(s, w->f2y, i, nhalf) += LAT(nhalf) / nlon / AC(nlat) * dt * pn;
  }
  for (int j = 2; j <= nlat - 1; j++)
    for (int i = 1; i <= nlon; i++) {
      const double g = A(s, an, i, j);
      A(s, aq, i, j) += A(s, s->old_m, i, j) * g;
      A(s, w->f2x, i, j) -= LEL(j) / AC(j) * dt * g;
      A(s, w->f2x, i - 1, j) += LEL(j) / AC(j) * dt * g;
      A(s, w->f2y, i, j) -= LAT(j) / AC(j) * dt * g;
      A(s, w->f2y, i, j - 1) += LAT(j - 1) / AC(j) * dt * g;
    }
  fill_halo_ad(s, w->f2y, KH, 0, 0, 1, 0);
  fill_halo_ad(s, w->f2x, KF, 1, 0, 0, 0);
  tl_zero(s, aqx);
  tl_zero(s, aqy);
  hflx_ad(s, w, s->mfx, s->mfy, s->qy, s->qx, w->f2x, w->f2y, aqy, aqx);
  fill_halo_ad(s, aqy, KF, 1, 1, 0, 0);
  fill_halo_ad(s, aqx, KF, 0, 0, 1, 1);
  tl_zero(s, w->f1x);
  tl_zero(s, w->f1y);
  ps = pn = 0;
  for (int i = 1; i <= nlon; i++) {
    A(s, aq, i, 1) += A(s, aqx, i, 1) + (1 + 0.5 * A(s, s->divy, i, 1) * dt) * A(s, aqy, i, 1);
    A(s, aq, i, nlat) += A(s, aqx, i, nlat) + (1 - 0.5 * A(s, s->divy, i, nlat) * dt) * A(s, aqy, i, nlat);
    ps += A(s, aqy, i, 1);
    pn += A(s, aqy, i, nlat);
  }
  for (int i = 1; i <= nlon; i++) {
    A(s, w->f1y, i, 1) -= 0.5 * dt * LAT(1) / nlon / AC(1) * ps;
    A(s, w->f1y, i, nhalf) += 0.5 * dt * LAT(nhalf) / nlon / AC(nlat) * pn;
  }
  for (int j = 2; j <= nlat - 1; j++)
    for (int i = 1; i <= nlon; i++) {
      const double gx = A(s, aqx, i, j), gy = A(s, aqy, i, j);
      A(s, aq, i, j) += (1 + 0.5 * A(s, s->divx, i, j) * dt) * gx + (1 + 0.5 * A(s, s->divy, i, j) * dt) * gy;
      A(s, w->f1x, i, j) -= 0.5 * dt * LEL(j) / AC(j) * gx;
      A(s, w->f1x, i - 1, j) += 0.5 * dt * LEL(j) / AC(j) * gx;
      A(s, w->f1y, i, j) -= 0.5 * dt * LAT(j) / AC(j) * gy;
      A(s, w->f1y, i, j - 1) += 0.5 * dt * LAT(j - 1) / AC(j) * gy;
    }
  fill_halo_ad(s, w->f1y, KH, 0, 0, 1, 1);
  fill_halo_ad(s, w->f1x, KF, 1, 1, 0, 0);
  hflx_ad(s, w, s->u, s->v, qo, qo, w->f1x, w->f1y, aq, aq);
  fill_halo_ad(s, aq, KF, 1, 1, 1, 1);
}

/* One adv_run step (as gmo_adv_run) that also carries, for every tracer,
 * a tangent dq (old level in, new level out) and an adjoint lam (new level
 * in, old level out); either may be NULL.  Both are [ntracers][nlat][ld]
 * interior arrays.  A step that does not advance leaves them unchanged. */
int gmo_adv_tlad_run(gmo_adv *s, int64_t time_step, double *dq, double *lam, int64_t ld) {
  if (s->err) return s->err;
  if (ld < s->nlon || s->cfg.flux_type != GMC_FLUX_PPM) return GMC_EINVAL;   /* TL/AD of the PPM flux only */
  if (!(time_step > 0 && time_step % s->cfg.nstep == 0)) return gmo_adv_run(s, time_step);
  accum_uv_cell(s);
  accum_mf_cell(s);
  if (s->err) return s->err;
  struct gmo_tlad *w = tlad_new(s);
  const size_t tstride = (size_t)s->nlat * ld;
  for (int l = 0; l < s->cfg.ntracers; l++) {
    advance_tracer(s, l, w->pre);
    if (dq) {
      tl_zero(s, w->dq);
      put(s, w->dq, s->nlat, dq + l * tstride, ld);
      tl_tracer(s, w, l);
      take(s, w->dn, s->nlat, dq + l * tstride, ld);
    }
    if (lam) {
      tl_zero(s, w->dn);
      put(s, w->dn, s->nlat, lam + l * tstride, ld);
      ad_tracer(s, w, l);
      take(s, w->dq, s->nlat, lam + l * tstride, ld);
    }
  }
  tlad_free(w);
  const int t = s->old;
  s->old = s->new_;
  s->new_ = t;
  memcpy(s->old_m, s->sm, (size_t)s->ncol * s->nrow * sizeof(double));
  return s->err;
}

/* a copy of the whole batch state */
gmo_adv *gmo_adv_clone(const gmo_adv *s) {
  gmo_adv *c = malloc(sizeof(*c));
  *c = *s;
  const size_t fb = (size_t)s->ncol * s->nrow * sizeof(double);
  const int mr = s->nlat + 2 * HY;
  c->mesh = malloc((size_t)GMC_MESH_NFIELDS * mr * sizeof(double));
  memcpy(c->mesh, s->mesh, (size_t)GMC_MESH_NFIELDS * mr * sizeof(double));
  c->wgt = malloc((s->nlat + 2) * sizeof(double));
  memcpy(c->wgt, s->wgt, (s->nlat + 2) * sizeof(double));
  double **f[] = {&c->su, &c->sv, &c->smfx, &c->smfy, &c->sm, &c->u, &c->v, &c->mfx, &c->mfy,
                  &c->old_m, &c->cflx, &c->cfly, &c->divx, &c->divy, &c->qx, &c->qy, &c->mlx,
                  &c->mly, &c->dmx, &c->dmy, &c->m6x, &c->m6y, &c->qmf_lon, &c->qmf_lat};
  for (size_t k = 0; k < sizeof(f) / sizeof(f[0]); k++) {
    const double *src = *f[k];
    *f[k] = malloc(fb);
    memcpy(*f[k], src, fb);
  }
  for (int lv = 0; lv < 2; lv++) {
    c->q[lv] = calloc(s->cfg.ntracers, sizeof(double *));
    for (int l = 0; l < s->cfg.ntracers; l++) {
      c->q[lv][l] = malloc(fb);
      memcpy(c->q[lv][l], s->q[lv][l], fb);
    }
  }
  return c;
}
