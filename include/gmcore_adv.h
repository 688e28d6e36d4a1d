/*
 * gmcore_adv.h -- C ABI of the FFSL tracer-advection path of the MI355X
 * (gfx950) library libgmcore_hip.so.
 *
 * The library is the GPU target of GMCORE's advection seam: the routine
 * `adv_run(block, itime)` (src/dynamics/adv/adv_mod.F90:152-312) with the
 * procedure pointer `adv_calc_tracer_hflx_cell => ffsl_calc_tracer_hflx_cell`
 * (adv_mod.F90:106-112, 122-130; src/dynamics/adv/ffsl_mod.F90:318-413) and
 * the wind accumulation of one `adv_batch_type` (adv_batch_mod.F90:19-69,
 * 218-403).  One gmc_adv object is one advection batch of one block: its
 * tracers q(:,:,1,l,1:2), the accumulated winds u, v, mfx, mfy, the CFL
 * numbers cflx, cfly, the split divergences divx, divy and old_m.
 *
 * Conventions are those of gmcore_hip.h: binary64, 0-based rows counted
 * from the south pole (full row r = reference j-1, half row r = half j-1),
 * interior-only host buffers with leading dimension `ld`, return 0 or a
 * GMC_E* code, one host thread and one HIP stream per object.
 */
#ifndef GMCORE_ADV_H
#define GMCORE_ADV_H

#include <stdint.h>
#include "gmcore_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* slope limiter of the PPM edge values (namelist limiter_type,
 * ffsl_mod.F90:131-140) */
enum gmc_limiter {
  GMC_LIMITER_MONO = 0,  /* 'mono', the default (slope_mono, ffsl_mod.F90:1183-1194) */
  GMC_LIMITER_PD = 1,    /* 'pd'   (slope_pd,   ffsl_mod.F90:1196-1205) */
  GMC_LIMITER_NONE = 2   /* 'none' (slope_simple, ffsl_mod.F90:1175-1181) */
};

/* horizontal flux of the FFSL tracer step (namelist ffsl_flux_type,
 * ffsl_mod.F90:120-129) */
enum gmc_flux_type {
  GMC_FLUX_PPM = 0,      /* 'ppm', the default (hflx_ppm_cell, ffsl_mod.F90:778-881)         */
  GMC_FLUX_VAN_LEER = 1  /* 'van_leer' (hflx_van_leer_cell, ffsl_mod.F90:556-616); no TL/AD */
};

typedef struct gmc_adv_config {
  double radius;            /* sphere radius, const_mod radius                     */
  double dt_adv;            /* batch%dt: advection time step in seconds            */
  int32_t nstep;            /* batch%nstep = dt_adv / dt_dyn (adv_batch_mod.F90:89) */
  int32_t ntracers;         /* tracers in the batch                                */
  int32_t lon_halo_width;   /* bound on |cflx| (adv_batch_mod.F90:284-287); below nlon/2 */
  int32_t limiter;          /* enum gmc_limiter                                    */
  int32_t lon_damp_passes;  /* lon_damp_on_cell(order 2) calls per step (3, adv_mod.F90:299-301) */
  int32_t flux_type;        /* enum gmc_flux_type                                  */
  const double *lon_damp_wgt; /* nlat weights wgt_full_lat of lon_damp_init
                                 (src/dynamics/damp/lon_damp_mod.F90:43-68); pole rows unused */
} gmc_adv_config;

typedef struct gmc_adv gmc_adv;

/* adv_batch_init + adv_allocate_tracers (adv_batch_mod.F90:73-145,
 * adv_mod.F90:334-406) for one cell-centred, non-dynamic batch.  All arrays
 * start at zero; time level old = 1, new = 2. */
int gmc_adv_create(const gmc_mesh *mesh, const gmc_adv_config *cfg, gmc_adv **out);
int gmc_adv_destroy(gmc_adv *adv);
const char *gmc_adv_last_error(const gmc_adv *adv);

/* q(:,:,1,tracer+1,level) <-> host, level 1 or 2, tracer 0-based.  Upload
 * also refreshes the periodic seam the reference keeps in the longitude
 * halo (fill_halo after set_ic, deform_test_mod.F90:78-124). */
int gmc_adv_upload_tracer(gmc_adv *adv, int32_t level, int32_t tracer, const double *host, int64_t ld);
int gmc_adv_download_tracer(gmc_adv *adv, int32_t level, int32_t tracer, double *host, int64_t ld);

/* Fields of block%state(itime) that adv_run reads (adv_mod.F90:420-455,
 * 228-297, 309): u_f (nlat rows), v_f (nlat-1 rows), mfx_lon (nlat rows),
 * mfy_lat (nlat-1 rows), m (nlat rows).  Any pointer may be NULL to keep the
 * field's current value. */
int gmc_adv_set_wind(gmc_adv *adv, const double *u_f, const double *v_f, const double *mfx_lon,
                     const double *mfy_lat, const double *m, int64_t ld);

/* adv_run(block, itime) (adv_mod.F90:152-312) at model step `time_step`:
 * accumulate the winds (adv_accum_wind, :420-455); when time_step > 0 and the
 * batch's alert rings (time_step a multiple of nstep, time_mod.F90:270-291)
 * advance every tracer from level old to new and swap; finally copy_old_m. */
int gmc_adv_run(gmc_adv *adv, int64_t time_step);

/* adv_run(block, itime) with block%state(itime) = state `level` of the
 * dynamics context `dyn` on the same grid (time_integrate, gmcore_mod.F90:588;
 * adv_mod.F90:152-312, adv_accum_wind :420-455): the batch accumulates that
 * state's u_f, v_f, mfx_lon and mfy_lat, advances its tracers when its
 * cadence rings with m of that state, and keeps m as old_m.  As in the
 * reference, mfx_lon and mfy_lat are those the level's last forward pass as a
 * star state left (after a pc2 step: the second stage's), so `dyn` must have
 * been stepped with gmc_keep_derived on (GMC_ESTATE otherwise).  The work is
 * ordered after the dynamics stream's work and the dynamics stream waits for
 * the copy of the fields; no device-host transfer. */
int gmc_adv_run_dyn(gmc_adv *adv, gmc_ctx *dyn, int32_t level, int64_t time_step);

/* What gmcore_run does for the batch before its time loop
 * (gmcore_mod.F90:117-131): copy_old_m and adv_accum_wind(block, old) on the
 * state the initial filters leave, whose mass fluxes no operator has formed
 * yet (zero).  `dyn` must be prepared; call it once, before the first
 * gmc_adv_run_dyn. */
int gmc_adv_init_dyn(gmc_adv *adv, gmc_ctx *dyn, int32_t level);

/* Prescribed-flow driver (gmcore_adv_driver.F90:84-92): upload once a wind
 * basis u_basis (nlat rows), v_basis (nlat-1 rows); gmc_adv_drive then does,
 * for n = 0 .. nsteps-1, set_uv with u_f = mfx_lon = u_basis * scale[n],
 * v_f = mfy_lat = v_basis * scale[n], m = 1 (the structure of
 * deform_case1_test_set_uv, deform_test_mod.F90:129-167), followed by
 * adv_run(time_step0 + n).  Device resident; one synchronisation at the end,
 * where the device guards (|cflx| > lon_halo_width, |cfly| > 1, a negative
 * tracer the fixer cannot fill) are checked. */
int gmc_adv_set_wind_basis(gmc_adv *adv, const double *u_basis, const double *v_basis, int64_t ld);
int gmc_adv_drive(gmc_adv *adv, int64_t time_step0, const double *scale, int32_t nsteps);

/* Tangent linear and adjoint of the prescribed-flow window with respect to
 * the tracers (no counterpart in the reference: the linearisation of adv_run,
 * adv_mod.F90:152-312, with ffsl_calc_tracer_hflx_cell, ffsl_mod.F90:318-413,
 * about a base trajectory, the winds held).  The step is piecewise linear in
 * q; the tangent follows the branches the base takes in the slope limiter,
 * the PPM edge clipping and the positivity fixer.
 *
 * gmc_adv_tl_drive runs gmc_adv_drive(time_step0, scale, nsteps) and maps dq
 * [ntracers][nlat][ld], the tangent of the old level at entry, to the tangent
 * of the old level at exit, in place.
 *
 * gmc_adv_ad_drive runs the same window and maps lam [ntracers][nlat][ld],
 * the adjoint of the old level at exit (for a cost J(q_N): dJ/dq_N), to the
 * adjoint of the old level at entry (dJ/dq_0), in place.  It keeps the slab
 * every ckpt_every steps, then, last segment first, recomputes the segment
 * from its checkpoint keeping every step's base state and runs the adjoint
 * steps back over it.  Memory: nsteps / ckpt_every + 1 checkpoints (copies of
 * the object's device slab) and two tape buffers (one when ckpt_every >=
 * nsteps) of one slot per advancing step of a segment (a step advances at time
 * steps k nstep, k > 0), each slot the step's base fields (m, the accumulated
 * winds, Courant numbers, split divergences, old_m, the tracers of both
 * levels, qx, qy, qp), kept between calls (the tangent keeps none: it runs
 * inside each forward step).  The object ends as after
 * gmc_adv_drive.  The adjoint adds with atomics: reproducible to rounding, not
 * bit for bit. */
int gmc_adv_tl_drive(gmc_adv *adv, int64_t time_step0, const double *scale, int32_t nsteps, double *dq, int64_t ld);
int gmc_adv_ad_drive(gmc_adv *adv, int64_t time_step0, const double *scale, int32_t nsteps, int32_t ckpt_every,
                     double *lam, int64_t ld);

/* Gradient of J = 1/2 sum over tracers and cells of (q_N - y)^2, q_N the
 * old level at the window's end, with respect to the old level at entry:
 * gmc_adv_ad_drive forced by q_N - y, formed on the device after the
 * window's forward (no extra forward).  y, grad: [ntracers][nlat][ld];
 * *cost = J (may be NULL). */
int gmc_adv_gradient(gmc_adv *adv, int64_t time_step0, const double *scale, int32_t nsteps, int32_t ckpt_every,
                     const double *y, double *grad, double *cost, int64_t ld);

/* Index (1 or 2) of the old time level after the swaps. */
int gmc_adv_old_level(const gmc_adv *adv, int32_t *old_level);

/* Tracer mass sum(q * area_cell) of every tracer on `level` (the driver's
 * diagnose, gmcore_adv_driver.F90:103-126, per tracer); out[ntracers]. */
int gmc_adv_mass(gmc_adv *adv, int32_t level, double *out);

/* Number of cells the horizontal positivity fixer (adv_mod.F90:236-258) has
 * filled since creation. */
int gmc_adv_fix_count(gmc_adv *adv, int64_t *count);

/* Wait for the object's stream; its HIP stream handle. */
int gmc_adv_synchronize(gmc_adv *adv);
int gmc_adv_stream(gmc_adv *adv, void **stream);

/* Per-kernel timing: runs gmc_adv_drive with a start and an end HIP event per
 * launch, both stamped by the dispatch itself (hipExtLaunchKernelGGL);
 * out_ms[k] = mean duration (ms) of kernel class k, gmc_adv_kernel_name(k). */
int gmc_adv_kernel_times(gmc_adv *adv, int64_t time_step0, const double *scale, int32_t nsteps,
                         double *out_ms, int32_t cap, int32_t *n);
const char *gmc_adv_kernel_name(int32_t k);

#ifdef __cplusplus
}
#endif

#endif /* GMCORE_ADV_H */
