/*
 * gmcore_baro.h -- hydrostatic column operators of GMCORE's baroclinic
 * dynamics on the MI355X (gfx950): the vertical-coordinate pressures, the
 * temperature, the geopotential on half levels, the surface-pressure
 * tendency, the vertical mass flux and its edge interpolations, the vertical
 * advection of momentum and the Lin (1997) pressure gradient.
 *
 * Each entry point is one reference subroutine of operators_mod / pgf_lin97_mod
 * (cited per function), run over every column of the block.  The horizontal
 * operators of a level (calc_mf, calc_ke, calc_vor, ...) are those of the
 * shallow-water library (gmcore_hip.h) applied per level.
 *
 * Conventions are those of gmcore_hip.h (rows counted from the south pole,
 * longitude fastest, interior only) plus a level index: full levels
 * k = 0 .. nlev-1 are the reference's k = 1 .. nlev counted from the model
 * top, half levels k = 0 .. nlev are the reference's 1 .. nlev+1 (half level
 * k is the top face of full level k).  A host buffer of a field holds all its
 * levels: element (i, r, k) at host[(k * rows + r) * ld + i], rows = nlat for
 * cell and U-point fields, nlat - 1 for V-point fields.
 */
#ifndef GMCORE_BARO_H
#define GMCORE_BARO_H

#include <stdint.h>
#include "gmcore_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The hybrid sigma-pressure coordinate (hybrid_coord_mod.F90:146-171):
 * ph_lev(k) = hyai(k) (p0 - ptop) + hybi(k) (phs - ptop) + ptop and
 * dphdt_lev(k) = hybi(k) dphs.  p0 is hybrid_coord_p0 (namelist_mod.F90:54),
 * ptop the coordinate's local_ptop (0 unless the template sets it,
 * hybrid_coord_mod.F90:27, 76-105).  A sigma coordinate is the case
 * hyai = 0. */
typedef struct gmc_vcoord {
  int32_t nlev;          /* num_full_lev, at least 2                    */
  int32_t reserved;      /* 0                                           */
  double p0;             /* hybrid_coord_p0                             */
  double ptop;           /* local_ptop                                  */
  const double *hyai;    /* nlev + 1 half-level coefficients            */
  const double *hybi;    /* nlev + 1                                    */
} gmc_vcoord;

/* Thermodynamic constants of const_mod (const_mod.F90:56-83). */
typedef struct gmc_thermo {
  double Rd;             /* J kg-1 K-1                                  */
  double cpd;            /* J kg-1 K-1; Rd_o_cpd = Rd / cpd             */
  double p0;             /* reference pressure of temperature()         */
} gmc_thermo;

/* state_type / tend_type / static fields of the column operators */
enum gmc_baro_field {
  GMB_PHS = 0,       /* state%phs_f       cells, 1 level                 */
  GMB_GZS,           /* static%gzs        cells, 1 level                 */
  GMB_DPHS,          /* tend%dphs         cells, 1 level                 */
  GMB_PT,            /* state%pt          cells, full levels             */
  GMB_PH_LEV,        /* state%ph_lev      cells, half levels             */
  GMB_PH_EXN_LEV,    /* state%ph_exn_lev  cells, half levels             */
  GMB_PH,            /* state%ph          cells, full levels             */
  GMB_T,             /* state%t           cells, full levels             */
  GMB_GZ_LEV,        /* state%gz_lev      cells, half levels             */
  GMB_GZ,            /* state%gz          cells, full levels             */
  GMB_DMFDLON,       /* tend%dmfdlon      cells, full levels             */
  GMB_DMFDLAT,       /* tend%dmfdlat      cells, full levels             */
  GMB_WE_LEV,        /* state%we_lev      cells, half levels             */
  GMB_WE_LEV_LON,    /* state%we_lev_lon  U points, half levels          */
  GMB_WE_LEV_LAT,    /* state%we_lev_lat  V points, half levels          */
  GMB_U_F,           /* state%u_f         U points, full levels          */
  GMB_V_F,           /* state%v_f         V points, full levels          */
  GMB_M_LON,         /* state%m_lon       U points, full levels          */
  GMB_M_LAT,         /* state%m_lat       V points, full levels          */
  GMB_WEDUDLEV,      /* tend%wedudlev     U points, full levels          */
  GMB_WEDVDLEV,      /* tend%wedvdlev     V points, full levels          */
  GMB_PGF_LON,       /* tend%pgf_lon      U points, full levels          */
  GMB_PGF_LAT,       /* tend%pgf_lat      V points, full levels          */
  GMB_CFLZ,          /* batch%cflz        cells, half levels             */
  GMB_PTF_LEV,       /* state%ptf_lev     cells, half levels             */
  GMB_DPTFDLEV,      /* tend%dptfdlev     cells, full levels             */
  GMB_U_LON,         /* state%u_lon       U points, full levels          */
  GMB_V_LAT,         /* state%v_lat       V points, full levels          */
  GMB_ADV_U,         /* adv_batch_mass%u     U points, full levels       */
  GMB_ADV_V,         /* adv_batch_mass%v     V points, full levels       */
  GMB_ADV_MFX,       /* adv_batch_mass%mfx   U points, full levels       */
  GMB_ADV_MFY,       /* adv_batch_mass%mfy   V points, full levels       */
  GMB_ADV_CFLX,      /* adv_batch_mass%cflx  U points, full levels       */
  GMB_ADV_CFLY,      /* adv_batch_mass%cfly  V points, full levels       */
  GMB_ADV_DIVX,      /* adv_batch_mass%divx  cells, full levels          */
  GMB_ADV_DIVY,      /* adv_batch_mass%divy  cells, full levels          */
  GMB_PTF_LON,       /* state%ptf_lon     U points, full levels          */
  GMB_PTF_LAT,       /* state%ptf_lat     V points, full levels          */
  GMB_DPTFDLON,      /* tend%dptfdlon     cells, full levels             */
  GMB_DPTFDLAT,      /* tend%dptfdlat     cells, full levels             */
  GMB_NFIELDS
};

typedef struct gmc_baro gmc_baro;

/* One block of nlev levels on the current device, every field zero. */
int gmc_baro_create(const gmc_mesh *mesh, const gmc_vcoord *vc, const gmc_thermo *th, gmc_baro **out);
int gmc_baro_destroy(gmc_baro *b);
const char *gmc_baro_last_error(const gmc_baro *b);

/* All levels of a field <-> host (layout above); ld >= nlon. */
int gmc_baro_set(gmc_baro *b, int32_t field, const double *host, int64_t ld);
int gmc_baro_get(gmc_baro *b, int32_t field, double *host, int64_t ld);

/* calc_ph (operators_mod.F90:136-168): ph_lev and ph_exn_lev = ph_lev **
 * Rd_o_cpd on the half levels from phs, ph on the full levels. */
int gmc_baro_calc_ph(gmc_baro *b);
/* calc_t (operators_mod.F90:170-190): t = temperature(pt, ph)
 * (formula_mod.F90:24-31), pt (ph / p0) ** Rd_o_cpd. */
int gmc_baro_calc_t(gmc_baro *b);
/* calc_gz_lev (operators_mod.F90:433-469): gz_lev on half levels 0 ..
 * nlev-1 from gzs, t and ph_lev (each level's sum from that level down, in
 * ascending order), gz on the full levels.  The bottom half level keeps its
 * value: gmcore_run sets it to gzs once (gmcore_mod.F90:109). */
int gmc_baro_calc_gz_lev(gmc_baro *b);
/* calc_dphsdt (operators_mod.F90:1120-1143): dphs = -sum over the levels
 * of dmfdlon + dmfdlat. */
int gmc_baro_calc_dphsdt(gmc_baro *b);
/* calc_we_lev (operators_mod.F90:192-232) without the advection batch's
 * accumulation: we_lev on the interior half levels from dphs and the
 * column's mass-flux divergence above each level, 0 at the top and the
 * bottom; we_lev_lon and we_lev_lat by the area-weighted edge interpolations
 * with no wind given (interp_mod.F90:360-369, :425-434; the U points of the
 * pole rows are not written). */
int gmc_baro_calc_we_lev(gmc_baro *b);
/* calc_wedudlev_wedvdlev (operators_mod.F90:1145-1216): the SB81 vertical
 * advection of u_f and v_f by we_lev_lon / we_lev_lat over m_lon / m_lat
 * (U points off the pole rows, every V point). */
int gmc_baro_calc_wedudlev_wedvdlev(gmc_baro *b);
/* pgf_lin97_run, hydrostatic branch (pgf_lin97_mod.F90:20-105): pgf_lon on
 * the U points off the pole rows, pgf_lat on every V point, from ph_exn_lev
 * and gz_lev. */
int gmc_baro_pgf_lin97(gmc_baro *b);

/* The vertical part of calc_grad_ptf (operators_mod.F90:1084-1098): pt's
 * two levels beyond the top and the bottom extrapolated linearly, the
 * vertical FFSL flux of pt (vflx_ppm_cell, ffsl_mod.F90:883-941, PPM with
 * the limiter GMC_LIMITER_* of include/gmcore_adv.h) through the interior
 * half levels with the Courant numbers cflz and the mass flux we_lev (the
 * advection batch's we), into ptf_lev, and dptfdlev = ptf_lev(k+1) -
 * ptf_lev(k).  ptf_lev's top and bottom faces are not written.  The call
 * only queues the work; a Courant number whose upwind cell falls outside the
 * column (the reference reads beyond its arrays there) makes the next
 * gmc_baro_synchronize return GMC_ENUMERIC. */
int gmc_baro_calc_ptf_lev(gmc_baro *b, int32_t limiter);

/* The horizontal part of calc_grad_ptf (operators_mod.F90:1029-1080): on
 * every level, ffsl_calc_tracer_hflx_cell (ffsl_mod.F90:318-413) of pt with
 * the mass batch's fields GMB_ADV_* and the time step dt: the inner fluxes of
 * pt with the winds adv_u / adv_v, the advective half-steps qx, qy (pole
 * rows from zonal sums), the outer fluxes of qy along x and qx along y with
 * the mass fluxes adv_mfx / adv_mfy into ptf_lon (the U points of the rows
 * off the poles) and ptf_lat; then dptfdlon on the rows off the poles and
 * dptfdlat, whose pole rows take the zonal sum of the adjacent ptf_lat row.
 * limiter: GMC_LIMITER_*; flux_type: GMC_FLUX_* (include/gmcore_adv.h);
 * lon_halo_width: the bound on |adv_cflx| (below nlon / 2).  The call only
 * queues the work; a Courant number beyond the bound (adv_batch_mod.F90:
 * 284-287) makes the next gmc_baro_synchronize return GMC_ENUMERIC. */
int gmc_baro_calc_ptf_hflx(gmc_baro *b, int32_t limiter, int32_t flux_type, int32_t lon_halo_width, double dt);
/* calc_grad_ptf (operators_mod.F90:1008-1118): gmc_baro_calc_ptf_hflx, then
 * gmc_baro_calc_ptf_lev with the same limiter (the vertical fluxes use the
 * batch's we and cflz, GMB_WE_LEV and GMB_CFLZ). */
int gmc_baro_calc_grad_ptf(gmc_baro *b, int32_t limiter, int32_t flux_type, int32_t lon_halo_width, double dt);

/* The table `small_filter2` of update_state's small_filter2 step (filter_init's
 * 'small_filter2', filter_types_mod.F90:124-143: the big filter's widths
 * times exp_two_values(filter_coef_e, filter_coef_d, lat0, 80, |lat|)),
 * copied to the device once, as filter_init builds it once per run; a later
 * call replaces it (after the queued work that reads it). */
int gmc_baro_set_small_filter2(gmc_baro *b, const gmc_filter *small_filter2);
/* The baroclinic branch of update_state's small_filter2 step
 * (time_schemes_mod.F90:213-221): filter_on_lon_edge / filter_on_lat_edge
 * (filter_mod.F90:93-163) with the table set above, in place on every level
 * of u_lon and v_lat; rows with at most one tap are left as they are.
 * GMC_EINVAL when no table is set. */
int gmc_baro_small_filter2(gmc_baro *b);

/* Wait for the block's queued work (GMC_ENUMERIC if a calc_ptf_lev or
 * calc_ptf_hflx since the last call met a Courant number out of range); its
 * stream (hipStream_t). */
int gmc_baro_synchronize(gmc_baro *b);
int gmc_baro_stream(gmc_baro *b, void **stream);

#ifdef __cplusplus
}
#endif

#endif /* GMCORE_BARO_H */
