/*
 * gmcore_hip.h -- C ABI of the MI355X (gfx950) dynamics library.
 *
 * The library is the GPU target of GMCORE's dynamics seam: the procedure
 * pointers `time_integrator`, `step` and `space_operators`
 * (src/dynamics/time_schemes_mod.F90:23-59, src/dynamics/gmcore_mod.F90:41,73)
 * and the routine `update_state` (time_schemes_mod.F90:120).  A Fortran host
 * binds these functions with ISO_C_BINDING `bind(C)` interfaces.
 *
 * Conventions
 *  - All floating point is IEEE binary64.
 *  - Row index r is 0-based and counts from the south: full-latitude row r is
 *    the reference's j = r + 1 (r = 0 is the south pole cell row, r = nlat-1
 *    the north pole); half-latitude row r is the reference's half j = r + 1
 *    (r = 0 .. nlat-2).
 *  - Host field buffers hold the interior only, longitude fastest:
 *    element (i, r) at host[r * ld + i], i = 0 .. nlon-1 being the
 *    reference's i = 1 .. nlon.  `ld` >= nlon.  This is the interior of the
 *    reference's column-major, halo-padded array (mesh_mod.F90:597-609) read
 *    one latitude row at a time.
 *  - Every function returns 0 on success and a positive GMC_E* code on
 *    failure; gmc_last_error() gives the message.  The reference aborts
 *    instead (log_error / process_stop, src/utils/process_mod.F90:54-68); the
 *    Fortran shim turns a non-zero code back into its log_error call.
 *  - A context belongs to one host thread and one HIP stream; it is not
 *    re-entrant.
 */
#ifndef GMCORE_HIP_H
#define GMCORE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
enum {
  GMC_OK = 0,
  GMC_EINVAL = 1,     /* bad argument (size, index, null pointer)            */
  GMC_EDEVICE = 2,    /* HIP runtime error or no usable gfx950 device         */
  GMC_ENOMEM = 3,     /* device allocation failed                            */
  GMC_ENUMERIC = 4,   /* a device guard tripped (NaN, CFL beyond bounds ...)  */
  GMC_ESTATE = 5      /* call out of order (e.g. step before upload)         */
};

/* ---- mesh metrics -------------------------------------------------------
 * Packed 1-D metric table: metrics[f * nlat + r] for field f, row r.  Field
 * names are those of mesh_type (src/dynamics/mesh_mod.F90:15-128).  Full-
 * latitude fields use r = 0..nlat-1, half-latitude fields r = 0..nlat-2.
 * The host computes them once (mesh_init_global, mesh_mod.F90:134-447). */
enum gmc_mesh_field {
  GMC_FULL_LAT = 0, GMC_HALF_LAT, GMC_FULL_LAT_DEG, GMC_HALF_LAT_DEG,
  GMC_FULL_COS_LAT, GMC_HALF_COS_LAT, GMC_FULL_SIN_LAT, GMC_HALF_SIN_LAT,
  GMC_AREA_CELL, GMC_AREA_LON, GMC_AREA_LON_WEST, GMC_AREA_LON_EAST,
  GMC_AREA_LON_NORTH, GMC_AREA_LON_SOUTH,
  GMC_AREA_LAT, GMC_AREA_LAT_WEST, GMC_AREA_LAT_EAST, GMC_AREA_LAT_NORTH,
  GMC_AREA_LAT_SOUTH, GMC_AREA_VTX, GMC_AREA_SUBCELL1, GMC_AREA_SUBCELL2,
  GMC_DE_LON, GMC_LE_LON, GMC_DE_LAT, GMC_LE_LAT,
  GMC_FULL_F, GMC_HALF_F,
  GMC_FULL_TANGENT_WGT1, GMC_FULL_TANGENT_WGT2,
  GMC_HALF_TANGENT_WGT1, GMC_HALF_TANGENT_WGT2,
  GMC_MESH_NFIELDS
};

typedef struct gmc_mesh {
  int32_t nlon;            /* num_full_lon (even)                           */
  int32_t nlat;            /* num_full_lat, poles included                  */
  double dlon;             /* mesh%dlon                                     */
  const double *metrics;   /* GMC_MESH_NFIELDS * nlat doubles              */
} gmc_mesh;

/* ---- zonal filter table -------------------------------------------------
 * One filter_type (src/dynamics/filter_types_mod.F90:15-27): per full row the
 * tap count ngrid_lon[r] and weights wgt_lon[r * max_taps + k]; per half row
 * ngrid_lat / wgt_lat likewise.  A row with ngrid <= 1 is left unfiltered
 * (filter_mod.F90:25-50). */
typedef struct gmc_filter {
  int32_t max_taps;
  const int32_t *ngrid_lon;  /* nlat     */
  const double *wgt_lon;     /* nlat     * max_taps */
  const int32_t *ngrid_lat;  /* nlat - 1 */
  const double *wgt_lat;     /* (nlat-1) * max_taps */
} gmc_filter;

/* scheme switches: namelist keys of &gmcore_control
 * (src/utils/namelist_mod.F90:10-122) that select code on the SW path */
enum { GMC_TIME_PC2 = 0, GMC_TIME_WRFRK3 = 1, GMC_TIME_EULER = 2 };
enum { GMC_PV_MIDPOINT = 0, GMC_PV_UPWIND = 1, GMC_PV_TVD = 2 };
/* tvd_limiter_type (tvd_mod.F90:27-37) */
enum { GMC_TVD_NONE = 0, GMC_TVD_VAN_LEER = 1, GMC_TVD_MC = 2 };

typedef struct gmc_config {
  double g;                  /* gravity, const_mod.F90:61                     */
  int32_t time_scheme;       /* GMC_TIME_*   (time_schemes_mod.F90:67-76)     */
  int32_t pv_scheme;         /* GMC_PV_*     (operators_mod.F90:56-71)        */
  int32_t upwind_order_pv;   /* 1 or 3       (operators_mod.F90:738-795)      */
  double upwind_wgt_pv;
  int32_t pv_pole_stokes;    /* 0/1          (operators_mod.F90:640-674)      */
  int32_t coriolis_scheme;   /* 1 or 2       (operators_mod.F90:846-906)      */
  int32_t use_div_damp;      /* 0/1          (damp_mod.F90:56-58)             */
  int32_t div_damp_order;    /* 2 or 4       (div_damp_mod.F90:27-139)        */
  double div_damp_coef2;
  gmc_filter big_filter;     /* block%big_filter    (block_mod.F90:65)        */
  gmc_filter small_filter1;  /* block%small_filter1 (block_mod.F90:66)        */
  double div_damp_coef4;     /* order-4 coefficient (namelist_mod.F90:108)    */
  double dt_dyn;             /* the dt operators_prepare_1 runs with
                                (gmcore_mod.F90:133); read by pv_scheme tvd  */
  int32_t tvd_limiter;       /* GMC_TVD_*, for pv_scheme GMC_PV_TVD           */
  int32_t reserved;          /* 0                                            */
} gmc_config;

/* ---- fields of state_type / tend_type that cross the boundary ----------- */
enum gmc_field {
  GMC_U_LON = 0,  /* state%u_lon  U grid, nlat rows (pole rows unused) */
  GMC_V_LAT,      /* state%v_lat  V grid, nlat-1 rows                  */
  GMC_GZ,         /* state%gz     cells,  nlat rows                    */
  GMC_U_F,        /* state%u_f    filtered u                           */
  GMC_V_F,        /* state%v_f    filtered v                           */
  GMC_GZ_F,       /* state%gz_f   filtered gz                          */
  GMC_GZS,        /* static%gzs   surface geopotential (level ignored) */
  GMC_DU,         /* tend%du      U grid                               */
  GMC_DV,         /* tend%dv      V grid                               */
  GMC_DGZ,        /* tend%dgz     cells                                */
  GMC_PV,         /* state%pv     vertices, nlat-1 rows                */
  GMC_KE,         /* state%ke     cells                                */
  GMC_DIV,        /* state%div    cells                                */
  GMC_M,          /* state%m      cells                                */
  GMC_NFIELDS
};

typedef struct gmc_ctx gmc_ctx;

/* Create a device context: uploads metrics and filter tables, allocates three
 * state levels and three tendency levels (block_mod.F90:81-97).  Fields start
 * at zero, as the reference's allocate_array does (allocator_mod.F90:29-54). */
int gmc_create(const gmc_mesh *mesh, const gmc_config *cfg, gmc_ctx **out);
int gmc_destroy(gmc_ctx *ctx);
const char *gmc_last_error(const gmc_ctx *ctx);

/* Copy one interior field between host and device state `level`
 * (1..3, block%state(level); tendencies use block%tend(level)). */
int gmc_upload(gmc_ctx *ctx, int32_t level, int32_t field, const double *host, int64_t ld);
int gmc_download(gmc_ctx *ctx, int32_t level, int32_t field, double *host, int64_t ld);

/* gmcore_run's set-up before the time loop (gmcore_mod.F90:112-122, :133):
 * big filter u->u_f, v->v_f, gz->gz_f and operators_prepare on `level`. */
int gmc_prepare(gmc_ctx *ctx, int32_t level);

/* `space_operators` (gmcore_mod.F90:343-575, SW branches) with the
 * reference's argument meaning: pass 0 = all, 1 = forward, 2 = backward. */
int gmc_space_operators(gmc_ctx *ctx, int32_t old_level, int32_t star_level,
                        int32_t new_level, int32_t tend1, int32_t tend2,
                        double dt, int32_t pass);

/* `update_state` (time_schemes_mod.F90:120-242, SW branch).  `what` is the
 * tendency's update flag set: 1 = update_gz, 2 = update_u and update_v. */
int gmc_update_state(gmc_ctx *ctx, int32_t tend, int32_t old_level,
                     int32_t new_level, double dt, int32_t what);

/* time_integrate (gmcore_mod.F90:577-591) for one step from level old_level
 * into new_level: `time_integrator` (pc2 / wrfrk3 / euler with step =>
 * step_forward_backward, time_schemes_mod.F90:102-118, 244-288) followed by
 * damp_run (damp_mod.F90:49-63).  Level 3 is the intermediate state.  The
 * caller swaps old and new afterwards, as time_advance does.  Uses the fused
 * stage kernels. */
int gmc_time_integrate(gmc_ctx *ctx, int32_t old_level, int32_t new_level, double dt);

/* `time_integrator` (time_schemes_mod.F90:244-288) with step =>
 * step_forward_backward, followed by damp_run (damp_mod.F90:49-63), then the
 * old/new swap of time_advance; repeated nsteps times.  Uses the fused
 * stage kernels, replayed from HIP graphs captured on first use for each
 * (old, new, dt): graphs of 8 steps while 8 or more remain, then graphs of
 * one step; device resident, nothing crosses PCIe. */
int gmc_step(gmc_ctx *ctx, double dt, int32_t nsteps);

/* Current time level index of `old` (1..3) after gmc_step's swaps. */
int gmc_current_level(const gmc_ctx *ctx, int32_t *old_level);

/* diagnose (gmcore_mod.F90:213-341, SW): out[0] = tm, out[1] = te,
 * out[2] = tpe, out[3] = tav.  m, m_lon and m_lat are those of the level's
 * current gz (what the backward pass's calc_m leaves).  The reference reads
 * the forward-pass fields (mfx_lon, mfy_lat, pv, pv_lon, pv_lat) as they are
 * stored in state(itime): after a pc2 or wrfrk3 step those of the second
 * stage's star state, after a euler step those of the step before.  When the
 * level holds them (gmc_prepare, gmc_space_operators, or steps taken with
 * gmc_keep_derived on), they are used as stored and te, tpe and tav are the
 * reference's values; otherwise they are formed from the level's own state
 * first (te, tpe and tav of the current fields). */
int gmc_diagnose(gmc_ctx *ctx, int32_t level, double out[4]);

/* gmc_diagnose without any allocation (no counterpart in the reference; for
 * callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same sums as gmc_diagnose,
 * bit for bit, written to d_out[0..3] in DEVICE memory, queued on the
 * context's stream (the call returns without waiting; gmc_synchronize or an
 * event on gmc_stream orders a read).  Its pole sums and row partial sums use
 * a static device array of the library, one per GPU, sized for at most 4097
 * latitude rows (about 230 KB); larger grids get GMC_EINVAL.  Every call on a
 * GPU shares that array: calls that overlap on one GPU (any contexts, any
 * streams) race and must be serialised by the caller. */
int gmc_diagnose_dev(gmc_ctx *ctx, int32_t level, double *d_out);

/* on = 1: every stage of gmc_step, gmc_time_integrate and gmc_kernel_times
 * also stores the star state's forward-pass fields (m_lon, m_lat, m_vtx,
 * mfx_lon, mfy_lat, mfx_lat, mfy_lon, u_lat, v_lon, pv, pv_lon, pv_lat) in
 * the star level, as operators_prepare_2 leaves them in state(star)
 * (operators_mod.F90:117-122), so that gmc_diagnose reads what the
 * reference's diagnose reads.  Costs about twelve field writes per stage; a
 * host loop turns it on for the step before a diagnostics print.  on = 0
 * (the default) skips those writes. */
int gmc_keep_derived(gmc_ctx *ctx, int32_t on);

/* Wait for all work queued on the context's stream. */
int gmc_synchronize(gmc_ctx *ctx);

/* Per-kernel timing of the fused step: runs nsteps steps exactly as gmc_step
 * does (the state advances), launched one by one on the context's stream,
 * each launch with a start and an end HIP event that the dispatch itself
 * stamps (hipExtLaunchKernelGGL), so the figure is the kernel's own duration
 * without the gap before it.  out_ms[k] receives the mean duration in
 * milliseconds of kernel class k (at most cap classes, *n set to the count);
 * gmc_kernel_name(k) names them. */
int gmc_kernel_times(gmc_ctx *ctx, double dt, int32_t nsteps, double *out_ms, int32_t cap, int32_t *n);
const char *gmc_kernel_name(int32_t k);

/* The HIP stream (hipStream_t) all work of the context is queued on. */
int gmc_stream(gmc_ctx *ctx, void **stream);

/* Device pointer of a field (for profilers and HIP-graph callers).  The
 * pointer stays valid until gmc_destroy.  Row pitch in doubles in *pitch. */
int gmc_field_ptr(gmc_ctx *ctx, int32_t level, int32_t field, double **dptr, int64_t *pitch);

/* Shape of the fused step's zonal-filter work list, built on first use (no
 * counterpart in the reference; for tests and profilers): out[0] = work
 * items of one stage, out[1] = latitude rows cut into more than one column
 * chunk, out[2] = most chunks of one row, out[3] = widest big filter (taps). */
int gmc_filter_plan(gmc_ctx *ctx, int32_t out[4]);

#ifdef __cplusplus
}
#endif

#endif /* GMCORE_HIP_H */
