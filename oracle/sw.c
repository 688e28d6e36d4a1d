/*
 * sw.c -- oracle restatement of GMCORE's shallow-water dynamics step.
 * TEST INFRASTRUCTURE ONLY (see gmc_oracle.h).
 *
 * Data layout as in the reference: every field is a halo-padded array
 * (1-hx:nlon+hx, 1-hy:nrow+hy), hy = 2 (gmcore_mod.F90:52-54, mesh_mod.F90:
 * 586-609), zero-initialised (allocator_mod.F90:29-54), with explicit
 * fill_halo calls (parallel_mod.F90:464-565 on one rank: periodic longitude,
 * cross-pole rows mirrored with a half-circle shift and no sign change,
 * halo_mod.F90:183-238).  Operators loop over the reference's index ranges.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "gmc_oracle.h"

#define HY 2
/* Row loops whose rows are independent run on all threads when the library
 * is built with -fopenmp (liboracle_omp.so); every element is still formed
 * by one thread in the same operation order, so both builds give the same
 * bits.  Reductions (zonal sums, diagnose) stay sequential. */
#ifdef _OPENMP
#define ROWS _Pragma("omp parallel for schedule(static)")
#else
#define ROWS
#endif
static const double EPS = 2.220446049250313e-16;  /* epsilon(1.0_r8), const_mod.F90:17 */

enum { KF = 0, KH = 1 };  /* full-lat / half-lat row grid */

/* state_type fields of the SW path (state_mod.F90:22-59) */
enum {
  S_U_LON, S_V_LAT, S_GZ, S_GZ_F, S_U_F, S_V_F, S_M, S_M_LON, S_M_LAT, S_M_VTX,
  S_MFX_LON, S_MFY_LAT, S_MFX_LAT, S_MFY_LON, S_U_LAT, S_V_LON, S_KE, S_DIV,
  S_VOR, S_PV, S_PV_LON, S_PV_LAT, S_DIV2, S_N
};
/* tend_type fields (tend_mod.F90:16-42) */
enum {
  T_DU, T_DV, T_DGZ, T_QHU, T_QHV, T_DKEDLON, T_DKEDLAT, T_DMFDLON, T_DMFDLAT,
  T_PGF_LON, T_PGF_LAT, T_N
};

typedef struct {
  int max_taps;
  int32_t *ngrid_lon, *ngrid_lat;  /* indexed by reference j (1-based) */
  double *wgt_lon, *wgt_lat;
} ofilter;

struct gmo_sw {
  int nlon, nlat, nhalf, hx, ncol, nrow_alloc;
  double dlon;
  double *mesh;          /* [GMC_MESH_NFIELDS][nlat + 2*HY], halo rows 0 */
  gmc_config cfg;
  ofilter big, small1;
  double *st[4][S_N];    /* state(1..3) */
  double *td[4][T_N];    /* tend(1..3)  */
  double *gzs;
  int old, new_;
};

/* mesh accessor, reference 1-based row j (halo rows read as 0) */
#define MS(s, f, j) ((s)->mesh[(f) * ((s)->nlat + 2 * HY) + (j) - 1 + HY])
/* field accessor, reference indices (i, j) */
#define A(s, a, i, j) ((a)[((j) - 1 + HY) * (s)->ncol + (i) - 1 + (s)->hx])

static double *falloc(const gmo_sw *s) { return calloc((size_t)s->ncol * s->nrow_alloc, sizeof(double)); }

/* ---------------------------------------------------------------------- */
/* fill_halo on one rank, parallel_mod.F90:464-565.  Longitude halos for rows
 * 1..nlat; then the pole exchange copies whole rows (halo columns included)
 * and reverses the interior columns by half a circle. */
static void fill_halo(gmo_sw *s, double *a, int kind, int west, int east, int south, int north) {
  const int nlon = s->nlon, hx = s->hx, nr = kind == KF ? s->nlat : s->nhalf;
  const int mx = nlon / 2;
  if (west)
    ROWS
    for (int j = 1; j <= s->nlat; j++)
      for (int i = 1 - hx; i <= 0; i++) A(s, a, i, j) = A(s, a, i + nlon, j);
  if (east)
    ROWS
    for (int j = 1; j <= s->nlat; j++)
      for (int i = nlon + 1; i <= nlon + hx; i++) A(s, a, i, j) = A(s, a, i - nlon, j);
  double *tmp = malloc(sizeof(double) * s->ncol * HY);
#define T(i, jj) tmp[((jj) - 1) * s->ncol + (i) - 1 + hx]
  if (south) {
    /* receive: row 1-HY+k <- row (KF: 2+k, KH: 1+k), k = 0..HY-1 */
    const int src0 = kind == KF ? 2 : 1;
    for (int k = 0; k < HY; k++)
      for (int i = 1 - hx; i <= nlon + hx; i++) A(s, a, i, 1 - HY + k) = A(s, a, i, src0 + k);
    for (int jj = 1; jj <= HY; jj++)
      for (int i = 1 - hx; i <= nlon + hx; i++) T(i, jj) = A(s, a, i, jj - HY);
    for (int jj = 1; jj <= HY; jj++) {
      const int j = 1 - jj;  /* assumed-shape row hy-jj+1 */
      for (int i = 1; i <= mx; i++) A(s, a, i, j) = T(i + mx, jj);
      for (int i = mx + 1; i <= nlon; i++) A(s, a, i, j) = T(i - mx, jj);
    }
  }
  if (north) {
    /* receive: row nr+1+k <- row nr-HY+k (KF) / nr-HY+1+k (KH) */
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
static void fill_all(gmo_sw *s, double *a, int kind) { fill_halo(s, a, kind, 1, 1, 1, 1); }

/* ---------------------------------------------------------------------- */
double gmo_zonal_sum(const double *x, int n, int stride) {
  double p[64];
  for (int l = 0; l < 64; l++) {
    double acc = 0.0;
    for (int i = l; i < n; i += 64) acc = acc + x[(size_t)i * stride];
    p[l] = acc;
  }
  for (int w = 1; w < 64; w *= 2)
    for (int l = 0; l < 64; l += 2 * w) p[l] = p[l] + p[l + w];
  return p[0];
}

/* zonal_sum over interior row j of a, after map f (parallel_mod.F90:741-764) */
static double row_sum(const gmo_sw *s, const double *a, int j, double sign, const double *scale) {
  double w[s->nlon];
  for (int i = 1; i <= s->nlon; i++) {
    double v = A(s, a, i, j);
    w[i - 1] = scale ? v * scale[0] : sign * v;
  }
  return gmo_zonal_sum(w, s->nlon, 1);
}

/* ---------------------------------------------------------------------- */
/* zonal filter of one field, filter_mod.F90:25-163.  y == NULL filters in
 * place; otherwise rows with n <= 1 are copied to y. */
static void filter_apply(gmo_sw *s, const ofilter *f, double *x, double *y, int kind) {
  const int nr = kind == KF ? s->nlat : s->nhalf;
  ROWS
  for (int j = 1; j <= nr; j++) {
    const int n = kind == KF ? f->ngrid_lon[j] : f->ngrid_lat[j];
    const double *w = (kind == KF ? f->wgt_lon : f->wgt_lat) + (size_t)j * f->max_taps;
    double tmp[s->ncol];
    if (n > 1) {
      const int hn = (n - 1) / 2;
      for (int i = 1; i <= s->nlon; i++) {
        double acc = 0.0;
        for (int k = 0; k < n; k++) acc = acc + w[k] * A(s, x, i - hn + k, j);
        tmp[i - 1 + s->hx] = acc;
      }
      /* y(:,j) = tmp copies the whole row; the halo part of tmp is undefined
       * in the reference and every caller refills it, so leave it as is */
      double *dst = y ? y : x;
      for (int i = 1; i <= s->nlon; i++) A(s, dst, i, j) = tmp[i - 1 + s->hx];
    } else if (y) {
      for (int i = 1 - s->hx; i <= s->nlon + s->hx; i++) A(s, y, i, j) = A(s, x, i, j);
    }
  }
}

/* ---------------------------------------------------------------------- */
/* calc_m SW branch + interpolations, operators_mod.F90:527-541,
 * interp_mod.F90:233-275, 438-463 */
static void calc_m(gmo_sw *s, double **S) {
  const double g = s->cfg.g;
  ROWS
  for (int j = 1; j <= s->nlat; j++)
    for (int i = 1; i <= s->nlon; i++) A(s, S[S_M], i, j) = (A(s, S[S_GZ], i, j) - A(s, s->gzs, i, j)) / g;
  fill_all(s, S[S_M], KF);
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, S[S_M_LON], i, j) = (A(s, S[S_M], i, j) + A(s, S[S_M], i + 1, j)) * 0.5;
  fill_all(s, S[S_M_LON], KF);
  ROWS
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, S[S_M_LAT], i, j) = (A(s, S[S_M], i, j) + A(s, S[S_M], i, j + 1)) * 0.5;
  fill_all(s, S[S_M_LAT], KH);
  ROWS
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, S[S_M_VTX], i, j) = ((A(s, S[S_M], i, j) + A(s, S[S_M], i + 1, j)) * MS(s, GMC_AREA_SUBCELL2, j) +
                                (A(s, S[S_M], i, j + 1) + A(s, S[S_M], i + 1, j + 1)) * MS(s, GMC_AREA_SUBCELL1, j + 1)) /
                               MS(s, GMC_AREA_VTX, j);
}

/* calc_mf, operators_mod.F90:545-610 (accum_uv_cell of the mass batch, whose
 * output the SW path never reads, is omitted) */
static void calc_mf(gmo_sw *s, double **S) {
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 0; i <= s->nlon; i++)
      A(s, S[S_MFX_LON], i, j) = A(s, S[S_M_LON], i, j) * A(s, S[S_U_F], i, j);
  ROWS
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon + 1; i++)
      A(s, S[S_MFY_LAT], i, j) = A(s, S[S_M_LAT], i, j) * A(s, S[S_V_F], i, j);
  ROWS
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++) {
      double v = MS(s, GMC_HALF_TANGENT_WGT1, j) * (A(s, S[S_MFX_LON], i - 1, j) + A(s, S[S_MFX_LON], i, j)) +
                 MS(s, GMC_HALF_TANGENT_WGT2, j) * (A(s, S[S_MFX_LON], i - 1, j + 1) + A(s, S[S_MFX_LON], i, j + 1));
      A(s, S[S_MFX_LAT], i, j) = v;
      A(s, S[S_U_LAT], i, j) = v / A(s, S[S_M_LAT], i, j);
    }
  fill_all(s, S[S_U_LAT], KH);
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++) {
      double v = MS(s, GMC_FULL_TANGENT_WGT1, j) * (A(s, S[S_MFY_LAT], i, j - 1) + A(s, S[S_MFY_LAT], i + 1, j - 1)) +
                 MS(s, GMC_FULL_TANGENT_WGT2, j) * (A(s, S[S_MFY_LAT], i, j) + A(s, S[S_MFY_LAT], i + 1, j));
      A(s, S[S_MFY_LON], i, j) = v;
      A(s, S[S_V_LON], i, j) = v / A(s, S[S_M_LON], i, j);
    }
  fill_all(s, S[S_V_LON], KF);
}

/* calc_ke with ke_scheme = 1 (forced for one level, namelist_mod.F90:228-235),
 * operators_mod.F90:234-258, 320-349 */
static void calc_ke(gmo_sw *s, double **S) {
  double *u = S[S_U_F], *v = S[S_V_F], *ke = S[S_KE];
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon + 1; i++)
      A(s, ke, i, j) = (MS(s, GMC_AREA_LON_WEST, j) * (A(s, u, i - 1, j) * A(s, u, i - 1, j)) +
                        MS(s, GMC_AREA_LON_EAST, j) * (A(s, u, i, j) * A(s, u, i, j)) +
                        MS(s, GMC_AREA_LAT_NORTH, j - 1) * (A(s, v, i, j - 1) * A(s, v, i, j - 1)) +
                        MS(s, GMC_AREA_LAT_SOUTH, j) * (A(s, v, i, j) * A(s, v, i, j))) /
                       MS(s, GMC_AREA_CELL, j);
  double sq[s->nlon];
  for (int i = 1; i <= s->nlon; i++) sq[i - 1] = A(s, v, i, 1) * A(s, v, i, 1);
  double pole = gmo_zonal_sum(sq, s->nlon, 1) / s->nlon;
  for (int i = 1; i <= s->nlon; i++) A(s, ke, i, 1) = pole;
  for (int i = 1; i <= s->nlon; i++) sq[i - 1] = A(s, v, i, s->nhalf) * A(s, v, i, s->nhalf);
  pole = gmo_zonal_sum(sq, s->nlon, 1) / s->nlon;
  for (int i = 1; i <= s->nlon; i++) A(s, ke, i, s->nlat) = pole;
}

/* calc_div, operators_mod.F90:354-431; div2 (the Laplacian of div on the
 * non-pole rows) when div_damp_order is 4 */
static void calc_div(gmo_sw *s, double **S) {
  double *u = S[S_U_F], *v = S[S_V_F], *div = S[S_DIV];
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, div, i, j) = ((A(s, u, i, j) * MS(s, GMC_LE_LON, j) - A(s, u, i - 1, j) * MS(s, GMC_LE_LON, j)) +
                         (A(s, v, i, j) * MS(s, GMC_LE_LAT, j) - A(s, v, i, j - 1) * MS(s, GMC_LE_LAT, j - 1))) /
                        MS(s, GMC_AREA_CELL, j);
  double pole = row_sum(s, v, 1, 1.0, NULL);
  pole = pole * MS(s, GMC_LE_LAT, 1) / s->nlon / MS(s, GMC_AREA_CELL, 1);
  for (int i = 1; i <= s->nlon; i++) A(s, div, i, 1) = pole;
  pole = row_sum(s, v, s->nhalf, -1.0, NULL);
  pole = pole * MS(s, GMC_LE_LAT, s->nhalf) / s->nlon / MS(s, GMC_AREA_CELL, s->nlat);
  for (int i = 1; i <= s->nlon; i++) A(s, div, i, s->nlat) = pole;
  if (s->cfg.div_damp_order != 4) {
    fill_halo(s, div, KF, 0, 1, 0, 1);
    return;
  }
  fill_halo(s, div, KF, 1, 1, 1, 1);
  double *div2 = S[S_DIV2];
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++) {
    const double dl = MS(s, GMC_DE_LON, j);
    for (int i = 1; i <= s->nlon; i++)
      A(s, div2, i, j) =
          (A(s, div, i + 1, j) - 2 * A(s, div, i, j) + A(s, div, i - 1, j)) / (dl * dl) +
          ((A(s, div, i, j + 1) - A(s, div, i, j)) * MS(s, GMC_HALF_COS_LAT, j) / MS(s, GMC_DE_LAT, j) -
           (A(s, div, i, j) - A(s, div, i, j - 1)) * MS(s, GMC_HALF_COS_LAT, j - 1) / MS(s, GMC_DE_LAT, j - 1)) /
              MS(s, GMC_LE_LON, j) / MS(s, GMC_FULL_COS_LAT, j);
  }
  fill_halo(s, div2, KF, 0, 1, 0, 1);
}

/* calc_vor + calc_pv, operators_mod.F90:612-703 */
static void calc_pv(gmo_sw *s, double **S) {
  double *u = S[S_U_F], *v = S[S_V_F], *vor = S[S_VOR];
  ROWS
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 0; i <= s->nlon; i++)
      A(s, vor, i, j) = (A(s, u, i, j) * MS(s, GMC_DE_LON, j) - A(s, u, i, j + 1) * MS(s, GMC_DE_LON, j + 1) +
                         A(s, v, i + 1, j) * MS(s, GMC_DE_LAT, j) - A(s, v, i, j) * MS(s, GMC_DE_LAT, j)) /
                        MS(s, GMC_AREA_VTX, j);
  if (s->cfg.pv_pole_stokes) {
    double w[s->nlon];
    int j = 1;
    for (int i = 1; i <= s->nlon; i++) w[i - 1] = -A(s, S[S_U_LAT], i, j) * MS(s, GMC_LE_LAT, j);
    double pole = gmo_zonal_sum(w, s->nlon, 1) / s->nlon / MS(s, GMC_AREA_CELL, j);
    for (int i = 1; i <= s->nlon; i++) A(s, vor, i, j) = pole;
    j = s->nhalf;
    for (int i = 1; i <= s->nlon; i++) w[i - 1] = A(s, S[S_U_LAT], i, j) * MS(s, GMC_LE_LAT, j);
    pole = gmo_zonal_sum(w, s->nlon, 1) / s->nlon / MS(s, GMC_AREA_CELL, j + 1);
    for (int i = 1; i <= s->nlon; i++) A(s, vor, i, j) = pole;
  }
  ROWS
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, S[S_PV], i, j) = (A(s, vor, i, j) + MS(s, GMC_HALF_F, j)) / A(s, S[S_M_VTX], i, j);
  fill_all(s, S[S_PV], KH);
}

/* upwind1 / upwind3, upwind_mod.F90:14-39 */
static double upwind1(double dir, double wgt, double f1, double f2) {
  return 0.5 * (f2 + f1) + -0.5 * (f2 - f1) * wgt * dir;
}
static double upwind3(double dir, double wgt, double f1, double f2, double f3, double f4) {
  const double c31 = 7.0 / 12.0, c32 = -1.0 / 12.0, c33 = 1.0 / 12.0;
  return c31 * (f3 + f2) + c32 * (f4 + f1) + c33 * (f4 - f1 - 3 * (f3 - f2)) * wgt * dir;
}

/* tvd with the flux limiters of tvd_mod.F90:42-100.  eps is the default-real
 * literal 1.0e-6 widened to r8 (tvd_mod.F90:51). */
static double tvd_limiter(int lim, double r) {
  if (lim == GMC_TVD_NONE) return 1;
  if (lim == GMC_TVD_VAN_LEER) return (r + fabs(r)) / (1 + fabs(r));
  return fmax(0.0, fmin(fmin(2.0 * r, 0.5 * (1 + r)), 2.0));
}
double gmo_tvd(int lim, double cfl, double fm1, double f, double fp1, double fp2) {
  const double eps = (double)1.0e-6f;
  const double r = cfl > 0 ? (f - fm1 + eps) / (fp1 - f + eps) : (fp2 - fp1 + eps) / (fp1 - f + eps);
  const double C = tvd_limiter(lim, r);
  return 0.5 * (fp1 + f - ((1 - C) * copysign(1.0, cfl) + cfl * C) * (fp1 - f));
}

/* interp_pv_upwind / interp_pv_midpoint / interp_pv_tvd,
 * operators_mod.F90:705-832; dt is the stage dt of operators_prepare */
static void interp_pv(gmo_sw *s, double **S, double dt) {
  double *pv = S[S_PV], *un = S[S_U_F], *vn = S[S_V_F], *ut = S[S_U_LAT], *vt = S[S_V_LON];
  const double wgt = s->cfg.upwind_wgt_pv;
  if (s->cfg.pv_scheme == GMC_PV_TVD) {
    const int lim = s->cfg.tvd_limiter;
    ROWS
    for (int j = 2; j <= s->nlat - 1; j++)
      for (int i = 1; i <= s->nlon; i++) {
        const double cfl = A(s, vt, i, j) * dt / MS(s, GMC_LE_LON, j);
        A(s, S[S_PV_LON], i, j) = gmo_tvd(lim, cfl, A(s, pv, i, j - 2), A(s, pv, i, j - 1), A(s, pv, i, j),
                                          A(s, pv, i, j + 1));
      }
    ROWS
    for (int j = 1; j <= s->nhalf; j++)
      for (int i = 1; i <= s->nlon; i++) {
        const double cfl = A(s, ut, i, j) * dt / MS(s, GMC_LE_LAT, j);
        A(s, S[S_PV_LAT], i, j) = gmo_tvd(lim, cfl, A(s, pv, i - 2, j), A(s, pv, i - 1, j), A(s, pv, i, j),
                                          A(s, pv, i + 1, j));
      }
    fill_halo(s, S[S_PV_LON], KF, 1, 1, 0, 1);
    fill_halo(s, S[S_PV_LAT], KH, 1, 1, 1, 0);
    return;
  }
  if (s->cfg.pv_scheme == GMC_PV_MIDPOINT) {
    ROWS
    for (int j = 1; j <= s->nhalf; j++)
      for (int i = 1; i <= s->nlon; i++)
        A(s, S[S_PV_LAT], i, j) = 0.5 * (A(s, pv, i - 1, j) + A(s, pv, i, j));
    fill_halo(s, S[S_PV_LAT], KH, 1, 1, 1, 0);
    ROWS
    for (int j = 2; j <= s->nlat - 1; j++)
      for (int i = 1; i <= s->nlon; i++)
        A(s, S[S_PV_LON], i, j) = 0.5 * (A(s, pv, i, j) + A(s, pv, i, j - 1));
    fill_halo(s, S[S_PV_LON], KF, 1, 1, 0, 1);
    return;
  }
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++) {
      double t = A(s, vt, i, j), n = A(s, un, i, j);
      double b = fabs(t) / (sqrt(n * n + t * t) + EPS);
      double up = s->cfg.upwind_order_pv == 1
          ? upwind1(copysign(1.0, t), wgt, A(s, pv, i, j - 1), A(s, pv, i, j))
          : upwind3(copysign(1.0, t), wgt, A(s, pv, i, j - 2), A(s, pv, i, j - 1), A(s, pv, i, j), A(s, pv, i, j + 1));
      A(s, S[S_PV_LON], i, j) = b * up + (1 - b) * 0.5 * (A(s, pv, i, j - 1) + A(s, pv, i, j));
    }
  ROWS
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++) {
      double t = A(s, ut, i, j), n = A(s, vn, i, j);
      double b = fabs(t) / (sqrt(t * t + n * n) + EPS);
      double up = s->cfg.upwind_order_pv == 1
          ? upwind1(copysign(1.0, t), wgt, A(s, pv, i - 1, j), A(s, pv, i, j))
          : upwind3(copysign(1.0, t), wgt, A(s, pv, i - 2, j), A(s, pv, i - 1, j), A(s, pv, i, j), A(s, pv, i + 1, j));
      A(s, S[S_PV_LAT], i, j) = b * up + (1 - b) * 0.5 * (A(s, pv, i - 1, j) + A(s, pv, i, j));
    }
  fill_halo(s, S[S_PV_LON], KF, 1, 1, 0, 1);
  fill_halo(s, S[S_PV_LAT], KH, 1, 1, 1, 0);
}

/* calc_grad_mf, operators_mod.F90:939-1006 */
static void calc_grad_mf(gmo_sw *s, double **S, double **T) {
  double *mfx = S[S_MFX_LON], *mfy = S[S_MFY_LAT];
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, T[T_DMFDLON], i, j) = (A(s, mfx, i, j) - A(s, mfx, i - 1, j)) * MS(s, GMC_LE_LON, j) / MS(s, GMC_AREA_CELL, j);
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, T[T_DMFDLAT], i, j) = (A(s, mfy, i, j) * MS(s, GMC_LE_LAT, j) - A(s, mfy, i, j - 1) * MS(s, GMC_LE_LAT, j - 1)) /
                                 MS(s, GMC_AREA_CELL, j);
  double pole = row_sum(s, mfy, 1, 1.0, NULL);
  pole = pole * MS(s, GMC_LE_LAT, 1) / s->nlon / MS(s, GMC_AREA_CELL, 1);
  for (int i = 1; i <= s->nlon; i++) A(s, T[T_DMFDLAT], i, 1) = pole;
  pole = row_sum(s, mfy, s->nhalf, -1.0, NULL);
  pole = pole * MS(s, GMC_LE_LAT, s->nhalf) / s->nlon / MS(s, GMC_AREA_CELL, s->nlat);
  for (int i = 1; i <= s->nlon; i++) A(s, T[T_DMFDLAT], i, s->nlat) = pole;
}

/* calc_coriolis, operators_mod.F90:846-906 */
static void calc_coriolis(gmo_sw *s, double **S, double **T) {
  double *mfx = S[S_MFX_LON], *mfy = S[S_MFY_LAT], *pl = S[S_PV_LON], *pt = S[S_PV_LAT];
  ROWS
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++) {
      double q;
      if (s->cfg.coriolis_scheme == 1) {
        double p = A(s, pt, i, j);
        q = (MS(s, GMC_HALF_TANGENT_WGT1, j) * (A(s, mfx, i - 1, j) * (p + A(s, pl, i - 1, j)) +
                                                A(s, mfx, i, j) * (p + A(s, pl, i, j))) +
             MS(s, GMC_HALF_TANGENT_WGT2, j) * (A(s, mfx, i - 1, j + 1) * (p + A(s, pl, i - 1, j + 1)) +
                                                A(s, mfx, i, j + 1) * (p + A(s, pl, i, j + 1)))) * 0.5;
      } else {
        q = A(s, S[S_MFX_LAT], i, j) * A(s, pt, i, j);
      }
      A(s, T[T_QHU], i, j) = q;
    }
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++) {
      double q;
      if (s->cfg.coriolis_scheme == 1) {
        double p = A(s, pl, i, j);
        q = (MS(s, GMC_FULL_TANGENT_WGT1, j) * (A(s, mfy, i, j - 1) * (p + A(s, pt, i, j - 1)) +
                                                A(s, mfy, i + 1, j - 1) * (p + A(s, pt, i + 1, j - 1))) +
             MS(s, GMC_FULL_TANGENT_WGT2, j) * (A(s, mfy, i, j) * (p + A(s, pt, i, j)) +
                                                A(s, mfy, i + 1, j) * (p + A(s, pt, i + 1, j)))) * 0.5;
      } else {
        q = A(s, S[S_MFY_LON], i, j) * A(s, pl, i, j);
      }
      A(s, T[T_QHV], i, j) = q;
    }
}

/* calc_grad_ke, operators_mod.F90:908-937 */
static void calc_grad_ke(gmo_sw *s, double **S, double **T) {
  double *ke = S[S_KE];
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, T[T_DKEDLON], i, j) = (A(s, ke, i + 1, j) - A(s, ke, i, j)) / MS(s, GMC_DE_LON, j);
  ROWS
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, T[T_DKEDLAT], i, j) = (A(s, ke, i, j + 1) - A(s, ke, i, j)) / MS(s, GMC_DE_LAT, j);
}

/* pgf_swm_run, pgf_swm_mod.F90:18-46 */
static void pgf_swm(gmo_sw *s, double **S, double **T) {
  double *gz = S[S_GZ_F];
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, T[T_PGF_LON], i, j) = (A(s, gz, i + 1, j) - A(s, gz, i, j)) / MS(s, GMC_DE_LON, j);
  ROWS
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++)
      A(s, T[T_PGF_LAT], i, j) = (A(s, gz, i, j + 1) - A(s, gz, i, j)) / MS(s, GMC_DE_LAT, j);
}

/* operators_prepare_2 forward pass, operators_mod.F90:117-122 */
static void prepare_forward(gmo_sw *s, double **S, double dt) {
  calc_mf(s, S);
  calc_ke(s, S);
  calc_div(s, S);
  calc_pv(s, S);
  interp_pv(s, S, dt);
}

/* ---------------------------------------------------------------------- */
int gmo_sw_space_operators(gmo_sw *s, int old_level, int star_level, int new_level,
                           int tend1, int tend2, double dt, int pass) {
  (void)old_level;
  if (star_level < 1 || star_level > 3 || new_level < 1 || new_level > 3 || tend1 < 1 || tend1 > 3 ||
      tend2 < 1 || tend2 > 3)
    return GMC_EINVAL;
  double **S = s->st[star_level], **N = s->st[new_level], **T1 = s->td[tend1], **T2 = s->td[tend2];
  const double g = s->cfg.g;
  if (pass == 0 || pass == 1) {
    /* operators_mod.F90:96-134; all_pass also runs calc_m first */
    if (pass == 0) calc_m(s, S);
    prepare_forward(s, S, dt);
    calc_grad_mf(s, S, T1);
    calc_coriolis(s, S, T1);
    calc_grad_ke(s, S, T1);
    if (pass == 0) pgf_swm(s, S, T1);
    ROWS
    for (int j = 2; j <= s->nlat - 1; j++)
      for (int i = 1; i <= s->nlon; i++)
        A(s, T1[T_DU], i, j) = pass == 0
            ? A(s, T1[T_QHV], i, j) - A(s, T1[T_PGF_LON], i, j) - A(s, T1[T_DKEDLON], i, j)
            : A(s, T1[T_QHV], i, j) - A(s, T1[T_DKEDLON], i, j);
    ROWS
    for (int j = 1; j <= s->nhalf; j++)
      for (int i = 1; i <= s->nlon; i++)
        A(s, T1[T_DV], i, j) = pass == 0
            ? -A(s, T1[T_QHU], i, j) - A(s, T1[T_PGF_LAT], i, j) - A(s, T1[T_DKEDLAT], i, j)
            : -A(s, T1[T_QHU], i, j) - A(s, T1[T_DKEDLAT], i, j);
    ROWS
    for (int j = 1; j <= s->nlat; j++)
      for (int i = 1; i <= s->nlon; i++)
        A(s, T1[T_DGZ], i, j) = -(A(s, T1[T_DMFDLON], i, j) + A(s, T1[T_DMFDLAT], i, j)) * g;
    return GMC_OK;
  }
  if (pass == 2) {
    /* gmcore_mod.F90:552-571 */
    calc_m(s, N);
    pgf_swm(s, N, T1);
    ROWS
    for (int j = 2; j <= s->nlat - 1; j++)
      for (int i = 1; i <= s->nlon; i++)
        A(s, T1[T_DU], i, j) = A(s, T2[T_DU], i, j) - A(s, T1[T_PGF_LON], i, j);
    ROWS
    for (int j = 1; j <= s->nhalf; j++)
      for (int i = 1; i <= s->nlon; i++)
        A(s, T1[T_DV], i, j) = A(s, T2[T_DV], i, j) - A(s, T1[T_PGF_LAT], i, j);
    return GMC_OK;
  }
  return GMC_EINVAL;
}

/* update_state, SW branch, time_schemes_mod.F90:178-239 */
int gmo_sw_update_state(gmo_sw *s, int tend, int old_level, int new_level, double dt, int what) {
  double **O = s->st[old_level], **N = s->st[new_level], **T = s->td[tend];
  if (what & 1) {
    ROWS
    for (int j = 1; j <= s->nlat; j++)
      for (int i = 1; i <= s->nlon; i++)
        A(s, N[S_GZ], i, j) = A(s, O[S_GZ], i, j) + dt * A(s, T[T_DGZ], i, j);
    fill_halo(s, N[S_GZ], KF, 1, 1, 0, 0);
    filter_apply(s, &s->big, N[S_GZ], N[S_GZ_F], KF);
    fill_all(s, N[S_GZ_F], KF);
    filter_apply(s, &s->small1, N[S_GZ], NULL, KF);
    fill_all(s, N[S_GZ], KF);
  }
  if (what & 2) {
    ROWS
    for (int j = 2; j <= s->nlat - 1; j++)
      for (int i = 1; i <= s->nlon; i++)
        A(s, N[S_U_LON], i, j) = A(s, O[S_U_LON], i, j) + dt * A(s, T[T_DU], i, j);
    ROWS
    for (int j = 1; j <= s->nhalf; j++)
      for (int i = 1; i <= s->nlon; i++)
        A(s, N[S_V_LAT], i, j) = A(s, O[S_V_LAT], i, j) + dt * A(s, T[T_DV], i, j);
    fill_all(s, N[S_U_LON], KF);
    fill_all(s, N[S_V_LAT], KH);
    const double wgt = 0.6;
    double *v = N[S_V_LAT];
    for (int i = 1 - s->hx; i <= s->nlon + s->hx; i++)
      A(s, v, i, 1) = (1 - wgt) * A(s, v, i, 1) + wgt * A(s, v, i, 2);
    for (int i = 1 - s->hx; i <= s->nlon + s->hx; i++)
      A(s, v, i, s->nhalf) = (1 - wgt) * A(s, v, i, s->nhalf) + wgt * A(s, v, i, s->nhalf - 1);
    fill_halo(s, v, KH, 0, 0, 1, 1);
    filter_apply(s, &s->big, N[S_U_LON], N[S_U_F], KF);
    filter_apply(s, &s->big, N[S_V_LAT], N[S_V_F], KH);
    fill_all(s, N[S_U_F], KF);
    fill_all(s, N[S_V_F], KH);
  }
  return GMC_OK;
}

/* div_damp_init / div_damp_run, orders 2 and 4, div_damp_mod.F90:27-139 (one
 * level, no baroclinic factor).  div comes from the filtered winds (calc_div
 * reads u_f, v_f); u_lon and v_lat are damped and u_f, v_f are not
 * refiltered. */
static void div_damp(gmo_sw *s, double **N) {
  calc_div(s, N);
  const int o4 = s->cfg.div_damp_order == 4;
  const double c2 = s->cfg.div_damp_coef2, c4 = s->cfg.div_damp_coef4;
  const double *d = o4 ? N[S_DIV2] : N[S_DIV];
  ROWS
  for (int j = 2; j <= s->nlat - 1; j++) {
    const double le = MS(s, GMC_LE_LON, j), de = MS(s, GMC_DE_LON, j);
    const double c = o4 ? c4 * (le * le) * (de * de) : c2 * le * de;
    for (int i = 1; i <= s->nlon; i++) {
      const double t = c * (A(s, d, i + 1, j) - A(s, d, i, j)) / de;
      A(s, N[S_U_LON], i, j) = o4 ? A(s, N[S_U_LON], i, j) - t : A(s, N[S_U_LON], i, j) + t;
    }
  }
  ROWS
  for (int j = 1; j <= s->nhalf; j++) {
    const double le = MS(s, GMC_LE_LAT, j), de = MS(s, GMC_DE_LAT, j);
    const double c = o4 ? c4 * (le * le) * (de * de) : c2 * le * de;
    for (int i = 1; i <= s->nlon; i++) {
      const double t = c * (A(s, d, i, j + 1) - A(s, d, i, j)) / de;
      A(s, N[S_V_LAT], i, j) = o4 ? A(s, N[S_V_LAT], i, j) - t : A(s, N[S_V_LAT], i, j) + t;
    }
  }
  fill_all(s, N[S_U_LON], KF);
  fill_all(s, N[S_V_LAT], KH);
}

/* damp_run's divergence damping alone on state(level) (damp_mod.F90:56-58) */
int gmo_sw_div_damp(gmo_sw *s, int level) {
  if (level < 1 || level > 3) return GMC_EINVAL;
  div_damp(s, s->st[level]);
  return GMC_OK;
}

/* step_forward_backward, time_schemes_mod.F90:102-118 */
static void step_fb(gmo_sw *s, int old, int star, int new_, int t1, int t2, double dt) {
  gmo_sw_space_operators(s, old, star, new_, t1, t2, dt, 1);
  gmo_sw_update_state(s, t1, old, new_, dt, 1);
  gmo_sw_space_operators(s, old, star, new_, t2, t1, dt, 2);
  gmo_sw_update_state(s, t2, old, new_, dt, 2);
}

int gmo_sw_step(gmo_sw *s, double dt, int nsteps) {
  for (int n = 0; n < nsteps; n++) {
    const int o = s->old, w = s->new_;
    switch (s->cfg.time_scheme) {
    case GMC_TIME_EULER:
      step_fb(s, o, o, w, o, w, dt);
      break;
    case GMC_TIME_WRFRK3:
      step_fb(s, o, o, w, o, w, dt / 3.0);
      step_fb(s, o, w, 3, o, w, dt / 2.0);
      step_fb(s, o, 3, w, o, w, dt);
      break;
    default: /* pc2, time_schemes_mod.F90:244-258 */
      step_fb(s, o, o, w, o, w, dt / 2.0);
      step_fb(s, o, w, 3, o, w, dt / 2.0);
      step_fb(s, o, 3, w, o, w, dt);
    }
    if (s->cfg.use_div_damp) div_damp(s, s->st[w]);
    s->old = w; s->new_ = o;  /* time_advance swaps the indices */
  }
  return GMC_OK;
}

int gmo_sw_current_level(const gmo_sw *s) { return s->old; }

int gmo_sw_prepare(gmo_sw *s, int level) {
  double **S = s->st[level];
  filter_apply(s, &s->big, S[S_U_LON], S[S_U_F], KF);
  fill_all(s, S[S_U_F], KF);
  filter_apply(s, &s->big, S[S_V_LAT], S[S_V_F], KH);
  fill_all(s, S[S_V_F], KH);
  filter_apply(s, &s->big, S[S_GZ], S[S_GZ_F], KF);
  fill_all(s, S[S_GZ_F], KF);
  /* operators_prepare_1 (operators_mod.F90:73-94), SW subset */
  calc_m(s, S);
  calc_mf(s, S);
  calc_ke(s, S);
  calc_pv(s, S);
  interp_pv(s, S, s->cfg.dt_dyn);
  calc_div(s, S);
  s->old = level;
  s->new_ = level == 1 ? 2 : 1;
  return GMC_OK;
}

/* diagnose, gmcore_mod.F90:213-341 (SW): the sums over the fields as they
 * are stored in state(level).  After a step those are m, m_lon, m_lat of the
 * backward pass's calc_m and mfx_lon, mfy_lat, pv, pv_lon, pv_lat of the
 * level's last forward pass as a star state (operators_mod.F90:117-122). */
static void diag_sums(gmo_sw *s, double **S, double out[4]) {
  double tm = 0, te_ke = 0, te_pe = 0, tav = 0, tpe = 0;
  const double g = s->cfg.g;
  for (int j = 1; j <= s->nlat; j++)
    for (int i = 1; i <= s->nlon; i++) tm = tm + A(s, S[S_M], i, j) * MS(s, GMC_AREA_CELL, j);
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++)
      te_ke = te_ke + A(s, S[S_MFX_LON], i, j) * 0.5 * A(s, S[S_U_LON], i, j) * MS(s, GMC_AREA_LON, j) * 2;
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++)
      te_ke = te_ke + A(s, S[S_MFY_LAT], i, j) * 0.5 * A(s, S[S_V_LAT], i, j) * MS(s, GMC_AREA_LAT, j) * 2;
  for (int j = 1; j <= s->nlat; j++)
    for (int i = 1; i <= s->nlon; i++) {
      double m = A(s, S[S_M], i, j);
      te_pe = te_pe + (m * m * g * 0.5 + m * A(s, s->gzs, i, j)) * MS(s, GMC_AREA_CELL, j);
    }
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++) tav = tav + A(s, S[S_PV], i, j) * MS(s, GMC_AREA_VTX, j);
  for (int j = 2; j <= s->nlat - 1; j++)
    for (int i = 1; i <= s->nlon; i++) {
      double p = A(s, S[S_PV_LON], i, j);
      tpe = tpe + A(s, S[S_M_LON], i, j) * (p * p) * 0.5 * MS(s, GMC_AREA_LON, j);
    }
  for (int j = 1; j <= s->nhalf; j++)
    for (int i = 1; i <= s->nlon; i++) {
      double p = A(s, S[S_PV_LAT], i, j);
      tpe = tpe + A(s, S[S_M_LAT], i, j) * (p * p) * 0.5 * MS(s, GMC_AREA_LAT, j);
    }
  out[0] = tm; out[1] = te_ke + 0.0 + te_pe; out[2] = tpe; out[3] = tav;
}

int gmo_sw_diagnose(gmo_sw *s, int level, double out[4]) {
  if (level < 1 || level > 3) return GMC_EINVAL;
  diag_sums(s, s->st[level], out);
  return GMC_OK;
}

/* the same sums after operators_prepare on the level's own state (fields of
 * one time; not what the reference reads after a step) */
int gmo_sw_diagnose_fresh(gmo_sw *s, int level, double out[4]) {
  if (level < 1 || level > 3) return GMC_EINVAL;
  double **S = s->st[level];
  calc_m(s, S); calc_mf(s, S); calc_ke(s, S); calc_pv(s, S); interp_pv(s, S, s->cfg.dt_dyn); calc_div(s, S);
  diag_sums(s, S, out);
  return GMC_OK;
}

/* ---------------------------------------------------------------------- */
static int copy_filter(gmo_sw *s, ofilter *o, const gmc_filter *f) {
  o->max_taps = f->max_taps;
  o->ngrid_lon = calloc(s->nlat + 2, sizeof(int32_t));
  o->ngrid_lat = calloc(s->nlat + 2, sizeof(int32_t));
  o->wgt_lon = calloc((size_t)(s->nlat + 2) * f->max_taps, sizeof(double));
  o->wgt_lat = calloc((size_t)(s->nlat + 2) * f->max_taps, sizeof(double));
  for (int r = 0; r < s->nlat; r++) {
    o->ngrid_lon[r + 1] = f->ngrid_lon[r];
    memcpy(o->wgt_lon + (size_t)(r + 1) * f->max_taps, f->wgt_lon + (size_t)r * f->max_taps, sizeof(double) * f->max_taps);
  }
  for (int r = 0; r < s->nhalf; r++) {
    o->ngrid_lat[r + 1] = f->ngrid_lat[r];
    memcpy(o->wgt_lat + (size_t)(r + 1) * f->max_taps, f->wgt_lat + (size_t)r * f->max_taps, sizeof(double) * f->max_taps);
  }
  int hw = 0;
  for (int r = 1; r <= s->nlat; r++) if ((o->ngrid_lon[r] - 1) / 2 > hw) hw = (o->ngrid_lon[r] - 1) / 2;
  for (int r = 1; r <= s->nhalf; r++) if ((o->ngrid_lat[r] - 1) / 2 > hw) hw = (o->ngrid_lat[r] - 1) / 2;
  return hw;
}

gmo_sw *gmo_sw_create(const gmc_mesh *mesh, const gmc_config *cfg) {
  if (!mesh || !cfg || mesh->nlon % 2 || mesh->nlat < 5) return NULL;
  gmo_sw *s = calloc(1, sizeof(gmo_sw));
  s->nlon = mesh->nlon; s->nlat = mesh->nlat; s->nhalf = mesh->nlat - 1; s->dlon = mesh->dlon;
  s->cfg = *cfg;
  /* process_create_blocks: lon halo = max(2, half-width of the big filter),
   * process_mod.F90:209-213 (the small filter is narrower) */
  int hw = copy_filter(s, &s->big, &cfg->big_filter);
  int hw2 = copy_filter(s, &s->small1, &cfg->small_filter1);
  s->hx = hw > 2 ? hw : 2;
  if (hw2 > s->hx) s->hx = hw2;
  if (s->hx > s->nlon) { gmo_sw_destroy(s); return NULL; }
  s->ncol = s->nlon + 2 * s->hx;
  s->nrow_alloc = s->nlat + 2 * HY;
  s->mesh = calloc((size_t)GMC_MESH_NFIELDS * (s->nlat + 2 * HY), sizeof(double));
  for (int f = 0; f < GMC_MESH_NFIELDS; f++)
    for (int r = 0; r < s->nlat; r++) MS(s, f, r + 1) = mesh->metrics[f * s->nlat + r];
  for (int l = 1; l <= 3; l++) {
    for (int f = 0; f < S_N; f++) s->st[l][f] = falloc(s);
    for (int f = 0; f < T_N; f++) s->td[l][f] = falloc(s);
  }
  s->gzs = falloc(s);
  s->old = 1; s->new_ = 2;
  return s;
}

void gmo_sw_destroy(gmo_sw *s) {
  if (!s) return;
  for (int l = 1; l <= 3; l++) {
    for (int f = 0; f < S_N; f++) free(s->st[l][f]);
    for (int f = 0; f < T_N; f++) free(s->td[l][f]);
  }
  ofilter *fs[2] = {&s->big, &s->small1};
  for (int k = 0; k < 2; k++) { free(fs[k]->ngrid_lon); free(fs[k]->ngrid_lat); free(fs[k]->wgt_lon); free(fs[k]->wgt_lat); }
  free(s->gzs); free(s->mesh); free(s);
}

int gmo_sw_halo_width(const gmo_sw *s) { return s->hx; }

/* map a gmcore_hip.h field id to (array, kind) */
static double *field_of(gmo_sw *s, int level, int field, int *kind) {
  if (level < 1 || level > 3) return NULL;
  double **S = s->st[level], **T = s->td[level];
  switch (field) {
  case GMC_U_LON: *kind = KF; return S[S_U_LON];
  case GMC_V_LAT: *kind = KH; return S[S_V_LAT];
  case GMC_GZ: *kind = KF; return S[S_GZ];
  case GMC_U_F: *kind = KF; return S[S_U_F];
  case GMC_V_F: *kind = KH; return S[S_V_F];
  case GMC_GZ_F: *kind = KF; return S[S_GZ_F];
  case GMC_GZS: *kind = KF; return s->gzs;
  case GMC_DU: *kind = KF; return T[T_DU];
  case GMC_DV: *kind = KH; return T[T_DV];
  case GMC_DGZ: *kind = KF; return T[T_DGZ];
  case GMC_PV: *kind = KH; return S[S_PV];
  case GMC_KE: *kind = KF; return S[S_KE];
  case GMC_DIV: *kind = KF; return S[S_DIV];
  case GMC_M: *kind = KF; return S[S_M];
  }
  return NULL;
}

int gmo_sw_set(gmo_sw *s, int level, int field, const double *host, int64_t ld) {
  int kind;
  double *a = field_of(s, level, field, &kind);
  if (!a || !host) return GMC_EINVAL;
  const int nr = kind == KF ? s->nlat : s->nhalf;
  for (int j = 1; j <= nr; j++)
    for (int i = 1; i <= s->nlon; i++) A(s, a, i, j) = host[(size_t)(j - 1) * ld + i - 1];
  /* set_ic routines fill the halo after writing the interior */
  fill_all(s, a, kind);
  return GMC_OK;
}

/* interior of any state (is_tend = 0, index S_*) or tendency (1, T_*) field */
int gmo_sw_get_internal(gmo_sw *s, int level, int is_tend, int index, double *host, int64_t ld) {
  static const int s_half[S_N] = {0, 1, 0, 0, 0, 1, 0, 0, 1, 1, 0, 1, 1, 0, 1, 0, 0, 0, 1, 1, 0, 1, 0};
  static const int t_half[T_N] = {0, 1, 0, 1, 0, 0, 1, 0, 0, 0, 1};
  if (level < 1 || level > 3 || !host) return GMC_EINVAL;
  if (index < 0 || index >= (is_tend ? T_N : S_N)) return GMC_EINVAL;
  double *a = is_tend ? s->td[level][index] : s->st[level][index];
  const int nr = (is_tend ? t_half[index] : s_half[index]) ? s->nhalf : s->nlat;
  for (int j = 1; j <= nr; j++)
    for (int i = 1; i <= s->nlon; i++) host[(size_t)(j - 1) * ld + i - 1] = A(s, a, i, j);
  return GMC_OK;
}

int gmo_sw_get(gmo_sw *s, int level, int field, double *host, int64_t ld) {
  int kind;
  double *a = field_of(s, level, field, &kind);
  if (!a || !host) return GMC_EINVAL;
  const int nr = kind == KF ? s->nlat : s->nhalf;
  for (int j = 1; j <= nr; j++)
    for (int i = 1; i <= s->nlon; i++) host[(size_t)(j - 1) * ld + i - 1] = A(s, a, i, j);
  return GMC_OK;
}
