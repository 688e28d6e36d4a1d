/*
 * baro.c -- CPU restatement of GMCORE's hydrostatic column operators
 * (TEST INFRASTRUCTURE ONLY: the checker of gmc_baro_* in the HIP library).
 *
 * Arrays use the host layout of include/gmcore_baro.h: element (i, r, k) of a
 * field at a[(k * rows + r) * nlon + i], rows = nlat for cells and U points,
 * nlat - 1 for V points; full levels 0 .. nlev-1, half levels 0 .. nlev from
 * the model top.  Each function follows the reference loop nest it cites,
 * with the reference's operation order; the build uses -ffp-contract=off and
 * libm's pow and log.
 */
#include <math.h>
#include <stddef.h>

#include "gmc_oracle.h"

#define MX(f, r) (metrics[(size_t)(f) * nlat + (r)])

/* calc_ph, operators_mod.F90:136-168, with hybrid_coord_calc_ph_lev
 * (hybrid_coord_mod.F90:155-162) and Rd_o_cpd = Rd / cpd (const_mod.F90:81) */
void gmo_baro_calc_ph(int nlon, int nlat, int nlev, const double *hyai, const double *hybi, double p0,
                      double ptop, double Rd, double cpd, const double *phs, double *ph_lev, double *ph_exn_lev,
                      double *ph) {
  const size_t L = (size_t)nlat * nlon;
  const double rd_o_cpd = Rd / cpd;
  for (int k = 0; k <= nlev; k++)
    for (size_t c = 0; c < L; c++) {
      ph_lev[k * L + c] = hyai[k] * (p0 - ptop) + hybi[k] * (phs[c] - ptop) + ptop;
      ph_exn_lev[k * L + c] = pow(ph_lev[k * L + c], rd_o_cpd);
    }
  for (int k = 0; k < nlev; k++)
    for (size_t c = 0; c < L; c++) ph[k * L + c] = 0.5 * (ph_lev[k * L + c] + ph_lev[(k + 1) * L + c]);
}

/* calc_t, operators_mod.F90:170-190; temperature(), formula_mod.F90:24-31 */
void gmo_baro_calc_t(int nlon, int nlat, int nlev, double Rd, double cpd, double p0, const double *pt,
                     const double *ph, double *t) {
  const size_t n = (size_t)nlev * nlat * nlon;
  const double rd_o_cpd = Rd / cpd;
  for (size_t c = 0; c < n; c++) t[c] = pt[c] * pow(ph[c] / p0, rd_o_cpd);
}

/* calc_gz_lev, operators_mod.F90:433-469: half levels 0 .. nlev-1 (the
 * bottom one keeps its value), then the full-level means */
void gmo_baro_calc_gz_lev(int nlon, int nlat, int nlev, double Rd, const double *gzs, const double *t,
                          const double *ph_lev, double *gz_lev, double *gz) {
  const size_t L = (size_t)nlat * nlon;
  for (int k = 0; k < nlev; k++)
    for (size_t c = 0; c < L; c++) {
      double dgz = 0.0;
      for (int l = k; l < nlev; l++) dgz = dgz + Rd * t[l * L + c] * log(ph_lev[(l + 1) * L + c] / ph_lev[l * L + c]);
      gz_lev[k * L + c] = gzs[c] + dgz;
    }
  for (int k = 0; k < nlev; k++)
    for (size_t c = 0; c < L; c++) gz[k * L + c] = 0.5 * (gz_lev[k * L + c] + gz_lev[(k + 1) * L + c]);
}

/* calc_dphsdt, operators_mod.F90:1120-1143 */
void gmo_baro_calc_dphsdt(int nlon, int nlat, int nlev, const double *dmfdlon, const double *dmfdlat,
                          double *dphs) {
  const size_t L = (size_t)nlat * nlon;
  for (size_t c = 0; c < L; c++) dphs[c] = 0.0;
  for (int k = 0; k < nlev; k++)
    for (size_t c = 0; c < L; c++) dphs[c] = dphs[c] - dmfdlon[k * L + c] - dmfdlat[k * L + c];
}

/* calc_we_lev, operators_mod.F90:192-232 (the advection batch's
 * accum_we_lev aside): every interior half level re-adds the divergence of
 * the levels above it from 0, as the reference does; then
 * interp_lev_edge_to_lev_lon_edge / _lat_edge with no wind
 * (interp_mod.F90:360-369, :425-434) */
void gmo_baro_calc_we_lev(int nlon, int nlat, int nlev, const double *metrics, const double *hybi,
                          const double *dmfdlon, const double *dmfdlat, const double *dphs, double *we_lev,
                          double *we_lev_lon, double *we_lev_lat) {
  const size_t L = (size_t)nlat * nlon, LV = (size_t)(nlat - 1) * nlon;
  for (int k = 1; k < nlev; k++)
    for (size_t c = 0; c < L; c++) {
      double mf = 0.0;
      for (int l = 0; l < k; l++) mf = mf + dmfdlon[l * L + c] + dmfdlat[l * L + c];
      we_lev[k * L + c] = -(hybi[k] * dphs[c]) - mf;
    }
  for (size_t c = 0; c < L; c++) {
    we_lev[c] = 0.0;
    we_lev[(size_t)nlev * L + c] = 0.0;
  }
  for (int k = 0; k <= nlev; k++) {
    const double *x = we_lev + k * L;
    for (int r = 1; r < nlat - 1; r++)
      for (int i = 0; i < nlon; i++) {
        const int ie = (i + 1) % nlon;
        we_lev_lon[k * L + (size_t)r * nlon + i] =
            (MX(GMC_AREA_LON_WEST, r) * x[(size_t)r * nlon + i] + MX(GMC_AREA_LON_EAST, r) * x[(size_t)r * nlon + ie]) /
            MX(GMC_AREA_LON, r);
      }
    for (int r = 0; r < nlat - 1; r++)
      for (int i = 0; i < nlon; i++)
        we_lev_lat[k * LV + (size_t)r * nlon + i] =
            (MX(GMC_AREA_LAT_NORTH, r) * x[(size_t)(r + 1) * nlon + i] + MX(GMC_AREA_LAT_SOUTH, r) * x[(size_t)r * nlon + i]) /
            MX(GMC_AREA_LAT, r);
  }
}

/* calc_wedudlev_wedvdlev, operators_mod.F90:1145-1216 (SB81) */
void gmo_baro_calc_wedudlev_wedvdlev(int nlon, int nlat, int nlev, const double *u, const double *v,
                                     const double *m_lon, const double *m_lat, const double *we_lev_lon,
                                     const double *we_lev_lat, double *wedudlev, double *wedvdlev) {
  const size_t L = (size_t)nlat * nlon, LV = (size_t)(nlat - 1) * nlon;
  for (int k = 1; k < nlev - 1; k++)
    for (int r = 1; r < nlat - 1; r++)
      for (int i = 0; i < nlon; i++) {
        const size_t c = (size_t)r * nlon + i;
        wedudlev[k * L + c] = (we_lev_lon[(k + 1) * L + c] * (u[(k + 1) * L + c] - u[k * L + c]) +
                               we_lev_lon[k * L + c] * (u[k * L + c] - u[(k - 1) * L + c])) /
                              m_lon[k * L + c] / 2.0;
      }
  for (int r = 1; r < nlat - 1; r++)
    for (int i = 0; i < nlon; i++) {
      const size_t c = (size_t)r * nlon + i, kb = nlev - 1;
      wedudlev[c] = (we_lev_lon[L + c] * (u[L + c] - u[c])) / m_lon[c] / 2.0;
      wedudlev[kb * L + c] = (we_lev_lon[kb * L + c] * (u[kb * L + c] - u[(kb - 1) * L + c])) / m_lon[kb * L + c] / 2.0;
    }
  for (int k = 1; k < nlev - 1; k++)
    for (size_t c = 0; c < LV; c++)
      wedvdlev[k * LV + c] = (we_lev_lat[(k + 1) * LV + c] * (v[(k + 1) * LV + c] - v[k * LV + c]) +
                              we_lev_lat[k * LV + c] * (v[k * LV + c] - v[(k - 1) * LV + c])) /
                             m_lat[k * LV + c] / 2.0;
  for (size_t c = 0; c < LV; c++) {
    const size_t kb = nlev - 1;
    wedvdlev[c] = (we_lev_lat[LV + c] * (v[LV + c] - v[c])) / m_lat[c] / 2.0;
    wedvdlev[kb * LV + c] = (we_lev_lat[kb * LV + c] * (v[kb * LV + c] - v[(kb - 1) * LV + c])) / m_lat[kb * LV + c] / 2.0;
  }
}

/* pgf_lin97_run, hydrostatic branch, pgf_lin97_mod.F90:59-105 */
void gmo_baro_pgf_lin97(int nlon, int nlat, int nlev, const double *metrics, const double *ph_exn_lev,
                        const double *gz_lev, double *pgf_lon, double *pgf_lat) {
  const size_t L = (size_t)nlat * nlon, LV = (size_t)(nlat - 1) * nlon;
  const double *e = ph_exn_lev, *z = gz_lev;
  for (int k = 0; k < nlev; k++) {
    const size_t t = k * L, b = (k + 1) * L;   /* the full level's top and bottom half levels */
    for (int r = 1; r < nlat - 1; r++)
      for (int i = 0; i < nlon; i++) {
        const size_t c = (size_t)r * nlon + i, ce = (size_t)r * nlon + (i + 1) % nlon;
        const double dph1 = e[b + ce] - e[t + c];
        const double dph2 = e[b + c] - e[t + ce];
        const double dgz1 = z[b + c] - z[t + ce];
        const double dgz2 = z[t + c] - z[b + ce];
        pgf_lon[k * L + c] = -(dph1 * dgz1 + dph2 * dgz2) / MX(GMC_DE_LON, r) / (dph1 + dph2);
      }
    for (int r = 0; r < nlat - 1; r++)
      for (int i = 0; i < nlon; i++) {
        const size_t c = (size_t)r * nlon + i, cn = (size_t)(r + 1) * nlon + i;
        const double dph1 = e[b + cn] - e[t + c];
        const double dph2 = e[b + c] - e[t + cn];
        const double dgz1 = z[b + c] - z[t + cn];
        const double dgz2 = z[t + c] - z[b + cn];
        pgf_lat[k * LV + c] = -(dph1 * dgz1 + dph2 * dgz2) / MX(GMC_DE_LAT, r) / (dph1 + dph2);
      }
  }
}

/* hybrid_coord_equal_eta (hybrid_coord_test_mod.F90:1330-1352): nlev + 1
 * half-level coefficients; etac is the default-real literal 0.2 widened. */
void gmo_baro_equal_eta(int nlev, double *hyai, double *hybi) {
  const double deta = 1.0 / nlev, etac = (double)0.2f;
  const double d3 = (1 - etac) * (1 - etac) * (1 - etac);
  const double c1 = 2 * (etac * etac) / d3;
  const double c2 = -etac * (4 + etac + etac * etac) / d3;
  const double c3 = 2 * (1 + etac + etac * etac) / d3;
  const double c4 = -(1 + etac) / d3;
  for (int k = 0; k <= nlev; k++) {
    const double eta = k * deta;
    hybi[k] = eta < etac ? 0.0 : c1 + c2 * eta + c3 * (eta * eta) + c4 * (eta * eta * eta);
    hyai[k] = eta - hybi[k];
  }
}

/* The vertical part of calc_grad_ptf (operators_mod.F90:1084-1098): the two
 * levels beyond the top and the bottom extrapolated linearly, then
 * vflx_ppm_cell (ffsl_mod.F90:883-941) through the interior half levels and
 * dptfdlev.  ptf_lev's top and bottom faces are left as given.  Returns the
 * number of faces whose upwind cell or whole cells fall outside the column
 * (their flux is left as given). */
int gmo_baro_calc_ptf_lev(int nlon, int nlat, int nlev, int limiter, const double *pt, const double *we,
                          const double *cflz, double *ptf_lev, double *dptfdlev) {
  const size_t L = (size_t)nlat * nlon;
  const int ne = nlev + 4;
  double ext[ne], ml[nlev], dm[nlev], m6[nlev];
  int bad = 0;
  for (size_t c = 0; c < L; c++) {
    for (int k = 0; k < nlev; k++) ext[k + 2] = pt[k * L + c];
    for (int k = -1; k >= -2; k--) ext[k + 2] = 2 * ext[k + 3] - ext[k + 4];
    for (int k = nlev; k <= nlev + 1; k++) ext[k + 2] = 2 * ext[k + 1] - ext[k];
    for (int k = 0; k < nlev; k++) {
      const double f[5] = {ext[k], ext[k + 1], ext[k + 2], ext[k + 3], ext[k + 4]};
      double out[3];
      gmo_ppm(limiter, f, out);
      ml[k] = out[0];
      dm[k] = out[1];
      m6[k] = out[2];
    }
    for (int k = 1; k < nlev; k++) {
      const double cf_all = cflz[k * L + c];
      const int ci = (int)cf_all;
      const double cf = cf_all - ci;
      if (cf_all > 0) {
        const int ku = k - ci - 1;
        if (ku < 0 || ku > nlev - 1 || k - ci < -2) { bad++; continue; }
        double sum = 0;
        for (int j = k - ci; j <= k - 1; j++) sum = sum + ext[j + 2];
        const double s1 = 1 - cf, s2 = 1;
        const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
        ptf_lev[k * L + c] = we[k * L + c] * (sum + ml[ku] * ds1 + 0.5 * dm[ku] * ds2 + m6[ku] * (ds2 / 2.0 - ds3 / 3.0)) / cf_all;
      } else if (cf_all < 0) {
        const int ku = k - ci;
        if (ku < 0 || ku > nlev - 1 || k - ci - 1 > nlev + 1) { bad++; continue; }
        double sum = 0;
        for (int j = k; j <= k - ci - 1; j++) sum = sum + ext[j + 2];
        const double s1 = 0, s2 = -cf;
        const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
        ptf_lev[k * L + c] = -we[k * L + c] * (sum + ml[ku] * ds1 + 0.5 * dm[ku] * ds2 + m6[ku] * (ds2 / 2.0 - ds3 / 3.0)) / cf_all;
      } else {
        ptf_lev[k * L + c] = 0;
      }
    }
    for (int k = 0; k < nlev; k++) dptfdlev[k * L + c] = ptf_lev[(k + 1) * L + c] - ptf_lev[k * L + c];
  }
  return bad;
}

/* update_state's small_filter2 step, baroclinic branch
 * (time_schemes_mod.F90:213-221): filter_on_lon_edge / filter_on_lat_edge
 * (filter_mod.F90:93-163) in place on every level, each output the sum of
 * its taps in order */
void gmo_baro_small_filter2(int nlon, int nlat, int nlev, const gmc_filter *f, double *u_lon, double *v_lat) {
  const size_t L = (size_t)nlat * nlon, LV = (size_t)(nlat - 1) * nlon;
  double tmp[nlon];
  for (int pass = 0; pass < 2; pass++) {
    const int rows = pass == 0 ? nlat : nlat - 1;
    const int32_t *ng = pass == 0 ? f->ngrid_lon : f->ngrid_lat;
    const double *wg = pass == 0 ? f->wgt_lon : f->wgt_lat;
    double *x = pass == 0 ? u_lon : v_lat;
    const size_t ls = pass == 0 ? L : LV;
    for (int k = 0; k < nlev; k++)
      for (int r = 0; r < rows; r++) {
        const int n = ng[r];
        if (n <= 1) continue;
        const int hn = (n - 1) / 2;
        const double *w = wg + (size_t)r * f->max_taps;
        double *row = x + k * ls + (size_t)r * nlon;
        for (int i = 0; i < nlon; i++) {
          double acc = 0.0;
          for (int t = 0; t < n; t++) acc = acc + w[t] * row[((i - hn + t) % nlon + nlon) % nlon];
          tmp[i] = acc;
        }
        for (int i = 0; i < nlon; i++) row[i] = tmp[i];
      }
  }
}

/* The horizontal part of calc_grad_ptf (operators_mod.F90:1029-1080): on
 * every level, ffsl_calc_tracer_hflx_cell of pt with the mass batch's
 * fields and the dynamics time step dt (adv.c's restatement) into ptf_lon
 * (U points of the rows off the poles) and ptf_lat (V points); dptfdlon on
 * the rows off the poles and dptfdlat, whose pole rows take the zonal sum of
 * the nearest ptf_lat row.  cfg: the batch's limiter, flux type and
 * lon_halo_width (ntracers >= 1, nstep >= 1, a damping table).  Returns 0,
 * or GMC_ENUMERIC when a cflx exceeds lon_halo_width (nothing written from
 * that level on). */
int gmo_baro_calc_ptf_hflx(const gmc_mesh *mesh, int nlev, const gmc_adv_config *cfg, double dt, const double *pt,
                           const double *u, const double *v, const double *mfx, const double *mfy,
                           const double *cflx, const double *cfly, const double *divx, const double *divy,
                           double *ptf_lon, double *ptf_lat, double *dptfdlon, double *dptfdlat) {
  const int nlon = mesh->nlon, nlat = mesh->nlat, nhalf = nlat - 1;
  const double *metrics = mesh->metrics;
  const size_t L = (size_t)nlat * nlon, Lv = (size_t)nhalf * nlon;
  gmo_adv *s = gmo_adv_create(mesh, cfg);
  if (!s) return GMC_EINVAL;
  double work[nlon];
  for (int k = 0; k < nlev; k++) {
    const int rc = gmo_adv_hflx_field(s, pt + k * L, u + k * L, v + k * Lv, mfx + k * L, mfy + k * Lv, cflx + k * L,
                                      cfly + k * Lv, divx + k * L, divy + k * L, dt, ptf_lon + k * L,
                                      ptf_lat + k * Lv, nlon);
    if (rc) {
      gmo_adv_destroy(s);
      return rc;
    }
    const double *fx = ptf_lon + k * L, *fy = ptf_lat + k * Lv;
    double *dx = dptfdlon + k * L, *dy = dptfdlat + k * L;
    for (int j = 1; j < nlat - 1; j++)
      for (int i = 0; i < nlon; i++) {
        const int iw = i == 0 ? nlon - 1 : i - 1;
        dx[j * nlon + i] = (fx[j * nlon + i] - fx[j * nlon + iw]) * MX(GMC_LE_LON, j) / MX(GMC_AREA_CELL, j);
      }
    for (int j = 1; j < nlat - 1; j++)
      for (int i = 0; i < nlon; i++)
        dy[j * nlon + i] = (fy[j * nlon + i] * MX(GMC_LE_LAT, j) - fy[(j - 1) * nlon + i] * MX(GMC_LE_LAT, j - 1)) /
                           MX(GMC_AREA_CELL, j);
    for (int i = 0; i < nlon; i++) work[i] = fy[i];
    double pole = gmo_zonal_sum(work, nlon, 1);
    pole = pole * MX(GMC_LE_LAT, 0) / nlon / MX(GMC_AREA_CELL, 0);
    for (int i = 0; i < nlon; i++) dy[i] = pole;
    for (int i = 0; i < nlon; i++) work[i] = -fy[(nhalf - 1) * nlon + i];
    pole = gmo_zonal_sum(work, nlon, 1);
    pole = pole * MX(GMC_LE_LAT, nhalf - 1) / nlon / MX(GMC_AREA_CELL, nlat - 1);
    for (int i = 0; i < nlon; i++) dy[(nlat - 1) * nlon + i] = pole;
  }
  gmo_adv_destroy(s);
  return 0;
}
