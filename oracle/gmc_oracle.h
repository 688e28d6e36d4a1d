/*
 * gmc_oracle.h -- CPU restatement of GMCORE's shallow-water dynamics path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker of the HIP library:
 * tests, the smoke check and the CPU-baseline leg of the benchmark call it.
 * No product code path may link or call it.
 *
 * It restates the reference algorithm with the reference's own data layout:
 * halo-padded arrays, explicit fill_halo calls (periodic in longitude,
 * mirrored across the poles), one routine per reference operator.  Each
 * function cites the reference file:line it follows.  Arithmetic is IEEE
 * binary64 in the source's left-to-right order; the build uses
 * -ffp-contract=off so no multiply-add is fused.  The mesh is computed with
 * __float128 where the reference uses real(16).
 */
#ifndef GMC_ORACLE_H
#define GMC_ORACLE_H

#include <stdint.h>
#include "../include/gmcore_hip.h"
#include "../include/gmcore_adv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* mesh_init_global (src/dynamics/mesh_mod.F90:134-447) into the packed
 * metric table of gmcore_hip.h (GMC_MESH_NFIELDS * nlat doubles). */
int gmo_mesh_init(int nlon, int nlat, double radius, double omega, double *metrics);

/* filter_init (src/dynamics/filter_types_mod.F90:49-147).  kind 0 =
 * big_filter, 1 = small_filter1, 2 = small_filter2.  First call with
 * wgt == NULL returns the tap counts and *max_taps; then call again with wgt
 * buffers sized nlat * max_taps and (nlat-1) * max_taps. */
int gmo_filter_init(int nlon, int nlat, const double *metrics, double dt,
                    double max_wave_speed, double max_cfl, double coef_a,
                    double coef_b, double coef_c, double coef_d, double coef_e, int kind,
                    int32_t *ngrid_lon, int32_t *ngrid_lat, int32_t *max_taps,
                    double *wgt_lon, double *wgt_lat);

/* Shallow-water model state of one block (block_mod.F90:27-45). */
typedef struct gmo_sw gmo_sw;

gmo_sw *gmo_sw_create(const gmc_mesh *mesh, const gmc_config *cfg);
void gmo_sw_destroy(gmo_sw *s);
int gmo_sw_halo_width(const gmo_sw *s);

/* interior copy in/out, same layout as gmc_upload/gmc_download */
int gmo_sw_set(gmo_sw *s, int level, int field, const double *host, int64_t ld);
int gmo_sw_get(gmo_sw *s, int level, int field, double *host, int64_t ld);

/* interior of any state_type field (is_tend = 0; index in the order u_lon,
 * v_lat, gz, gz_f, u_f, v_f, m, m_lon, m_lat, m_vtx, mfx_lon, mfy_lat,
 * mfx_lat, mfy_lon, u_lat, v_lon, ke, div, vor, pv, pv_lon, pv_lat) or
 * tend_type field (is_tend = 1; du, dv, dgz, qhu, qhv, dkedlon, dkedlat,
 * dmfdlon, dmfdlat, pgf_lon, pgf_lat) */
int gmo_sw_get_internal(gmo_sw *s, int level, int is_tend, int index, double *host, int64_t ld);

/* gmcore_run's initial filters + operators_prepare_1 (gmcore_mod.F90:112-133) */
int gmo_sw_prepare(gmo_sw *s, int level);
/* space_operators (gmcore_mod.F90:343-575), SW branches */
int gmo_sw_space_operators(gmo_sw *s, int old_level, int star_level, int new_level,
                           int tend1, int tend2, double dt, int pass);
/* update_state (time_schemes_mod.F90:120-242); what: 1 = gz, 2 = u and v */
int gmo_sw_update_state(gmo_sw *s, int tend, int old_level, int new_level,
                        double dt, int what);
/* time_integrator + damp_run + swap, nsteps times */
int gmo_sw_step(gmo_sw *s, double dt, int nsteps);
int gmo_sw_current_level(const gmo_sw *s);
/* diagnose (gmcore_mod.F90:213-341) over the fields stored in state(level) */
int gmo_sw_diagnose(gmo_sw *s, int level, double out[4]);
/* the same sums after operators_prepare on state(level) */
int gmo_sw_diagnose_fresh(gmo_sw *s, int level, double out[4]);
/* damp_run's divergence damping alone on state(level) (order 2 or 4) */
int gmo_sw_div_damp(gmo_sw *s, int level);
/* tvd (tvd_mod.F90:42-61) with limiter GMC_TVD_* */
double gmo_tvd(int lim, double cfl, double fm1, double f, double fp1, double fp2);

/* Pole sum in the fixed order the library uses: 64 strided partial sums
 * (partial p adds x[p], x[p+64], ... in ascending order) reduced by a
 * pairwise tree over p.  The reference's zonal_sum (parallel_mod.F90:741-764)
 * leaves the order to the compiler and MPI. */
double gmo_zonal_sum(const double *x, int n, int stride);

/* ---- FFSL tracer advection (adv.c) ------------------------------------- */

/* lon_damp_init (src/dynamics/damp/lon_damp_mod.F90:43-68): wgt[nlat] */
int gmo_lon_damp_weights(int nlat, const double *metrics, double *wgt);

/* One cell-centred advection batch (adv_batch_mod.F90:19-69) */
typedef struct gmo_adv gmo_adv;
gmo_adv *gmo_adv_create(const gmc_mesh *mesh, const gmc_adv_config *cfg);
void gmo_adv_destroy(gmo_adv *s);
int gmo_adv_set_tracer(gmo_adv *s, int level, int tracer, const double *host, int64_t ld);
int gmo_adv_get_tracer(gmo_adv *s, int level, int tracer, double *host, int64_t ld);
int gmo_adv_set_wind(gmo_adv *s, const double *u, const double *v, const double *mfx,
                     const double *mfy, const double *m, int64_t ld);
/* adv_run (adv_mod.F90:152-312) at model step time_step */
int gmo_adv_run(gmo_adv *s, int64_t time_step);
int gmo_adv_old_level(const gmo_adv *s);
int64_t gmo_adv_fix_count(const gmo_adv *s);
double gmo_adv_mass(gmo_adv *s, int level, int tracer);
/* interior of a batch array: 0 u, 1 v, 2 mfx, 3 mfy, 4 cflx, 5 cfly, 6 divx,
 * 7 divy, 8 qx, 9 qy, 10 qmf_lon, 11 qmf_lat, 12 old_m */
int gmo_adv_get_internal(gmo_adv *s, int index, double *host, int64_t ld);
int gmo_adv_hflx_field(gmo_adv *s, const double *q, const double *u, const double *v, const double *mfx,
                       const double *mfy, const double *cflx, const double *cfly, const double *divx,
                       const double *divy, double dt, double *qmfx, double *qmfy, int64_t ld);
/* tangent linear and adjoint of one adv_run step with respect to the
 * tracers (winds held): dq [ntracers][nlat][ld] old -> new level, lam new ->
 * old level; either may be NULL.  Advances the base like gmo_adv_run. */
int gmo_adv_tlad_run(gmo_adv *s, int64_t time_step, double *dq, double *lam, int64_t ld);
gmo_adv *gmo_adv_clone(const gmo_adv *s);
/* d (fl, df, f6) / d f[0..4] of ppm at the five values f */
void gmo_ppm_jac(int limiter, const double f[5], double J[3][5]);
/* ppm (ffsl_mod.F90:1147-1173) of f[0..4]: out = fl, df, f6 */
void gmo_ppm(int limiter, const double f[5], double out[3]);

/* Hydrostatic column operators (baro.c), arrays in the host layout of
 * include/gmcore_baro.h without padding (ld = nlon). */
void gmo_baro_calc_ph(int nlon, int nlat, int nlev, const double *hyai, const double *hybi, double p0,
                      double ptop, double Rd, double cpd, const double *phs, double *ph_lev, double *ph_exn_lev,
                      double *ph);
void gmo_baro_calc_t(int nlon, int nlat, int nlev, double Rd, double cpd, double p0, const double *pt,
                     const double *ph, double *t);
void gmo_baro_calc_gz_lev(int nlon, int nlat, int nlev, double Rd, const double *gzs, const double *t,
                          const double *ph_lev, double *gz_lev, double *gz);
void gmo_baro_calc_dphsdt(int nlon, int nlat, int nlev, const double *dmfdlon, const double *dmfdlat,
                          double *dphs);
void gmo_baro_calc_we_lev(int nlon, int nlat, int nlev, const double *metrics, const double *hybi,
                          const double *dmfdlon, const double *dmfdlat, const double *dphs, double *we_lev,
                          double *we_lev_lon, double *we_lev_lat);
void gmo_baro_calc_wedudlev_wedvdlev(int nlon, int nlat, int nlev, const double *u, const double *v,
                                     const double *m_lon, const double *m_lat, const double *we_lev_lon,
                                     const double *we_lev_lat, double *wedudlev, double *wedvdlev);
void gmo_baro_pgf_lin97(int nlon, int nlat, int nlev, const double *metrics, const double *ph_exn_lev,
                        const double *gz_lev, double *pgf_lon, double *pgf_lat);
/* the vertical part of calc_grad_ptf: ptf_lev (interior half levels) and
 * dptfdlev; returns the number of faces whose Courant number reaches
 * outside the column */
int gmo_baro_calc_ptf_lev(int nlon, int nlat, int nlev, int limiter, const double *pt, const double *we,
                          const double *cflz, double *ptf_lev, double *dptfdlev);
/* the horizontal part of calc_grad_ptf: ptf_lon, ptf_lat, dptfdlon, dptfdlat
 * of pt on every level with the mass batch's fields; 0 or GMC_ENUMERIC */
int gmo_baro_calc_ptf_hflx(const gmc_mesh *mesh, int nlev, const gmc_adv_config *cfg, double dt, const double *pt,
                           const double *u, const double *v, const double *mfx, const double *mfy,
                           const double *cflx, const double *cfly, const double *divx, const double *divy,
                           double *ptf_lon, double *ptf_lat, double *dptfdlon, double *dptfdlat);
/* small_filter2 of update_state's baroclinic branch, in place on u_lon and v_lat */
void gmo_baro_small_filter2(int nlon, int nlat, int nlev, const gmc_filter *f, double *u_lon, double *v_lat);
/* hybrid_coord_equal_eta (hybrid_coord_test_mod.F90:1330-1352) */
void gmo_baro_equal_eta(int nlev, double *hyai, double *hybi);

#ifdef __cplusplus
}
#endif

#endif
