/*
 * gmcore_io.h -- history (h0) and restart (r0) files of the shallow-water
 * path in the netCDF classic format (64-bit offset variant, "CDF-2").
 *
 * The reference writes them through its fiona library over netCDF
 * (history_setup_h0_swm / history_write_h0_swm, src/utils/history_mod.F90:
 * 125-150, 398-444; restart_write / restart_read, src/utils/restart_mod.F90:
 * 54-251).  fiona and netCDF are not part of this build, so the library
 * writes and reads the classic format itself: the same dimensions (time
 * unlimited, lon, lat, ilon, ilat), variables, dimension order, units and
 * long names.  Variables whose dtype the reference leaves to fiona's default
 * (tm and zs of h0) are written as 4-byte floats, the rest as doubles.
 *
 * Host arrays hold the interior, longitude fastest (element (i, r) at
 * a[r * nlon + i]), full-latitude fields nlat rows, half-latitude fields
 * nlat - 1 rows, as in include/gmcore_hip.h.  Files of the netCDF-4 (HDF5)
 * format are not read (GMC_EINVAL).
 */
#ifndef GMCORE_IO_H
#define GMCORE_IO_H

#include <stdint.h>
#include "gmcore_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One time record of h0 (history_write_h0_swm, history_mod.F90:398-444). */
typedef struct gmc_h0_record {
  int32_t nlon, nlat;
  const double *lon_deg;    /* nlon      full_lon_deg */
  const double *lat_deg;    /* nlat      full_lat_deg */
  const double *ilon_deg;   /* nlon      half_lon_deg */
  const double *ilat_deg;   /* nlat - 1  half_lat_deg */
  double time_value;        /* the record's time coordinate                    */
  const char *time_units;   /* units attribute of time, e.g. "days since ..."   */
  double dt;                /* global attribute time_step_size                  */
  double tm, te, tpe;       /* diagnose totals (gmcore_mod.F90:213-341)         */
  const double *u;          /* state%u_lon   (ilon, lat): nlat rows             */
  const double *v;          /* state%v_lat   (lon, ilat): nlat - 1 rows         */
  const double *z;          /* state%gz / g  (lon, lat)                         */
  const double *pv;         /* state%pv      (ilon, ilat): nlat - 1 rows        */
  const double *div;        /* state%div     (lon, lat)                         */
  const double *zs;         /* static%gzs / g (lon, lat), written with the first record */
} gmc_h0_record;

/* new_file = 1 creates path (history_setup_h0_swm's definitions) and writes
 * the coordinates, zs and record 0; new_file = 0 appends one record to a
 * file this function created (the coordinates and zs must match). */
int gmc_h0_write(const char *path, const gmc_h0_record *rec, int32_t new_file);

/* restart_write (restart_mod.F90:54-149), SW branch: one record of u, v,
 * gz and gzs with the coordinates and the attributes time_step_size and
 * restart_interval. */
int gmc_r0_write(const char *path, int32_t nlon, int32_t nlat, const double *lon_deg, const double *lat_deg,
                 const double *ilon_deg, const double *ilat_deg, double time_value, const char *time_units,
                 double dt_dyn, const char *restart_interval, const double *u, const double *v,
                 const double *gz, const double *gzs);

/* restart_read (restart_mod.F90:151-249), SW branch: u, v, gz, gzs and the
 * time value of record 0 of a classic-format file with the r0 layout on an
 * nlon x nlat grid. */
int gmc_r0_read(const char *path, int32_t nlon, int32_t nlat, double *u, double *v, double *gz, double *gzs,
                double *time_value);

/* The same from state `level` of a device context: h0 takes u, v, gz,
 * gzs and pv of the level and tm, te, tpe from gmc_diagnose (so pv and the
 * totals follow gmc_diagnose's rule for the level's derived fields); div is
 * that of the level's filtered winds (calc_div).  The fields are staged
 * through pinned host memory on the context's stream. */
int gmc_history_h0(gmc_ctx *ctx, const char *path, int32_t level, double time_value, const char *time_units,
                   double dt, int32_t new_file);
int gmc_restart_write(gmc_ctx *ctx, const char *path, int32_t level, double time_value, const char *time_units,
                      double dt_dyn, const char *restart_interval);
/* Reads u, v, gz and gzs into state `level` (the static gzs of the context);
 * the caller then runs gmc_prepare(level), as gmcore_run does after
 * restart_read. */
int gmc_restart_read(gmc_ctx *ctx, const char *path, int32_t level, double *time_value);

#ifdef __cplusplus
}
#endif

#endif /* GMCORE_IO_H */
