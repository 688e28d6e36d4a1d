// gmc_adv.hip -- FFSL tracer advection of one cell-centred advection batch on
// gfx950: the C ABI of include/gmcore_adv.h.
//
// Per step (adv_run, src/dynamics/adv/adv_mod.F90:152-312):
//   k_adv_accum   wind accumulation (adv_batch_mod.F90:218-252, 379-403)
//   k_adv_cfl     RK4 CFL numbers and split divergences (:254-373), only on
//                 the calls whose accumulation window closes
//   k_adv_inner   per (row, tracer): inner PPM fluxes of q along x (row in
//                 LDS) and y (the two V rows either side), then the advective
//                 half-steps qx, qy (ffsl_mod.F90:350-407)
//   k_adv_outer   per (row, tracer): outer PPM fluxes of qy along x and qx
//                 along y (ffsl_mod.F90:410), the flux-form update, /m, the
//                 sequential positivity fixer and the longitude damping of
//                 the row (adv_mod.F90:183-301)
// One workgroup owns one latitude row of one tracer, so the row-sequential
// parts (the fixer, the Jacobi damping) stay inside LDS and the only
// cross-row data are the qx, qy fields between the two tracer kernels.
//
// Layout in HBM: dense row-major [rows][pitch] doubles per field, no halo;
// tracer l of level k at q + ((k-1) * ntracers + l) * nlat * pitch.
// Longitude neighbours wrap; rows beyond a pole are the mirrored rows with a
// half-circle shift (the reference's fill_halo on one rank).  Every formula
// keeps the reference's operation order; the build uses -ffp-contract=off.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gmc_internal.hpp"
#include "gmc_point.hpp"
#include "../../include/gmcore_adv.h"

namespace gmc {
namespace {

constexpr int NA = 256;  // threads per workgroup of the row kernels
enum { F_CFLX = 1, F_CFLY = 2, F_NEG = 4 };
enum { KA_ACCUM, KA_CFL, KA_INNER, KA_OUTER, KA_NCLASS };
const char *const kAdvNames[KA_NCLASS] = {"k_adv_accum", "k_adv_cfl", "k_adv_inner", "k_adv_outer"};

struct AdvGrid {
  int nlon, nlat, nhalf, half;
  int64_t pitch;
  int hw;        // lon_halo_width: bound on |cflx|
  double dlon, radius, dt;
};

__device__ __forceinline__ int wrap_any(int i, int n) {
  const int r = i % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fsign(double a, double b) { return copysign(fabs(a), b); }

// slope_simple / slope_mono / slope_pd (ffsl_mod.F90:1175-1205)
__device__ __forceinline__ double slope(int lim, double fm1, double f, double fp1) {
  const double df = (fp1 - fm1) * 0.5;
  if (lim == GMC_LIMITER_NONE) return df;
  if (lim == GMC_LIMITER_PD) return fsign(dmin(fabs(df), 2 * f), df);
  const double df_min = 2 * (f - dmin(dmin(fm1, f), fp1));
  const double df_max = 2 * (dmax(dmax(fm1, f), fp1) - f);
  return fsign(dmin(dmin(fabs(df), df_min), df_max), df);
}

// ppm (ffsl_mod.F90:1147-1173)
struct Ppm { double fl, df, f6; };
__device__ __forceinline__ Ppm ppm(int lim, double fm2, double fm1, double f, double fp1, double fp2) {
  const double dfl = slope(lim, fm2, fm1, f);
  const double df = slope(lim, fm1, f, fp1);
  const double dfr = slope(lim, f, fp1, fp2);
  double fl = 0.5 * (fm1 + f) + (dfl - df) / 6.0;
  double fr = 0.5 * (fp1 + f) + (df - dfr) / 6.0;
  fl = f - fsign(dmin(fabs(df), fabs(fl - f)), df);
  fr = f + fsign(dmin(fabs(df), fabs(fr - f)), df);
  return Ppm{fl, fr - fl, 6 * f - 3 * (fl + fr)};
}

// The integral of the PPM parabola over [s1, s2] of the upwind cell, plus
// the whole cells crossed along x (hflx_ppm_cell, ffsl_mod.F90:831-877)
__device__ __forceinline__ double ppm_part_x(double s1, double s2, double ml, double dm, double m6,
                                             double sum) {
  const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
  return sum + ml * ds1 + 0.5 * dm * ds2 + m6 * (ds2 / 2.0 - ds3 / 3.0);
}
__device__ __forceinline__ double ppm_part_y(double s1, double s2, double ml, double dm, double m6) {
  const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
  return ml * ds1 + 0.5 * dm * ds2 + m6 * (ds2 / 2.0 - ds3 / 3.0);
}

// x-flux through U point i of a row whose PPM coefficients are in LDS
// (hflx_ppm_cell :831-853).  |int(cflx)| is clamped to the halo bound the
// accumulation step has already checked, so every index stays within one
// wrap of the row.
__device__ __forceinline__ double xflux(const AdvGrid &g, double c, double u, int i, const double *mx,
                                        const double *ml, const double *dm, const double *m6) {
  int ci = (int)c;
  const double cf = c - ci;
  ci = ci > g.hw ? g.hw : (ci < -g.hw ? -g.hw : ci);
  if (c > 0) {
    const int iu = wrap(i - ci, g.nlon);
    double sum = 0;
    for (int k = i + 1 - ci; k <= i; k++) sum = sum + mx[wrap(k, g.nlon)];
    return u * ppm_part_x(1 - cf, 1, ml[iu], dm[iu], m6[iu], sum) / c;
  } else if (c < 0) {
    const int iu = wrap(i - ci + 1, g.nlon);
    double sum = 0;
    for (int k = i + 1; k <= i - ci; k++) sum = sum + mx[wrap(k, g.nlon)];
    return -u * ppm_part_x(0, -cf, ml[iu], dm[iu], m6[iu], sum) / c;
  }
  return 0;
}

// y-flux through V point (i, h) of the field my (hflx_ppm_cell :855-877),
// with the PPM coefficients of the upwind full row computed here from the
// five cells of column i around it (rows through the poles mirrored).
__device__ __forceinline__ double yflux(const AdvGrid &g, int lim, double c, double v, int h, int i,
                                        const double *my) {
  if (c == 0) return 0;
  const int ju = c > 0 ? h : h + 1;
  double f[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    int r = ju - 2 + k, ii = i;
    if (r < 0) { r = -r; ii = wrap(i + g.half, g.nlon); }
    else if (r > g.nlat - 1) { r = 2 * (g.nlat - 1) - r; ii = wrap(i + g.half, g.nlon); }
    f[k] = my[r * g.pitch + ii];
  }
  const Ppm p = ppm(lim, f[0], f[1], f[2], f[3], f[4]);
  if (c > 0) return v * ppm_part_y(1 - c, 1, p.fl, p.df, p.f6) / c;
  return -v * ppm_part_y(0, -c, p.fl, p.df, p.f6) / c;
}

// PPM coefficients of a whole row held in LDS (ppm along x, :822-827)
__device__ __forceinline__ void ppm_row(const AdvGrid &g, int lim, const double *row, double *ml,
                                        double *dm, double *m6) {
  const int n = g.nlon;
  for (int i = threadIdx.x; i < n; i += NA) {
    const Ppm p = ppm(lim, row[wrap(i - 2, n)], row[wrap(i - 1, n)], row[i], row[wrap(i + 1, n)],
                      row[wrap(i + 2, n)]);
    ml[i] = p.fl;
    dm[i] = p.df;
    m6[i] = p.f6;
  }
}

// zonal sum of an LDS row in the library's fixed order, broadcast
__device__ __forceinline__ double block_zsum(const AdvGrid &g, const double *x, double *bcast) {
  __syncthreads();
  if (threadIdx.x < kWave) {
    const double s = wave_zonal_sum(g.nlon, threadIdx.x, [&](int i) { return x[i]; });
    if (threadIdx.x == 0) *bcast = s;
  }
  __syncthreads();
  return *bcast;
}

// ---------------------------------------------------------------------------
// Wind accumulation (adv_batch_mod.F90:238-251, 389-401).  mode 0: store,
// 1: average of nstep + 1 samples, 2: add.  With basis != null the sample is
// basis * scale (u_f = mfx_lon, v_f = mfy_lat of set_uv); otherwise the
// sample fields src_* are read.
struct AccumPtrs {
  double *u, *v, *mfx, *mfy;
  const double *su, *sv, *smfx, *smfy;
  const double *ub, *vb;
};

__global__ void __launch_bounds__(NA) k_adv_accum(AdvGrid g, AccumPtrs p, int mode, int nstep, double scale) {
  const int r = blockIdx.x;
  for (int i = threadIdx.x; i < g.nlon; i += NA) {
    const int64_t k = r * g.pitch + i;
    for (int f = 0; f < 4; f++) {
      if (r >= (f & 1 ? g.nhalf : g.nlat)) continue;
      double *acc = f == 0 ? p.u : f == 1 ? p.v : f == 2 ? p.mfx : p.mfy;
      double x;
      if (p.ub) {
        x = (f & 1 ? p.vb[k] : p.ub[k]) * scale;
      } else {
        const double *s = f == 0 ? p.su : f == 1 ? p.sv : f == 2 ? p.smfx : p.smfy;
        x = s[k];
      }
      if (mode == 0) acc[k] = x;
      else if (mode == 1) acc[k] = (acc[k] + x) / (double)(nstep + 1);
      else acc[k] = acc[k] + x;
    }
  }
}

// CFL numbers and split divergences (adv_batch_mod.F90:254-373); one
// workgroup per row index r: cflx and divx, divy of full row r, cfly of half
// row r; the pole rows' divy from the zonal sum of the adjacent V row.
struct CflPtrs {
  const double *u, *v;
  double *cflx, *cfly, *divx, *divy;
  int *flags;
};

__global__ void __launch_bounds__(NA) k_adv_cfl(AdvGrid g, Mesh M, CflPtrs p) {
  __shared__ double sv[4096 + 1];
  const int r = blockIdx.x, n = g.nlon;
  const double R = g.radius, dt = g.dt;
  int bad = 0;
  if (r > 0 && r < g.nlat - 1) {
    const double dx = M(GMC_DE_LON, r), cl = M(GMC_FULL_COS_LAT, r);
    const double *u = p.u + r * g.pitch;
    for (int i = threadIdx.x; i < n; i += NA) {
      // 1-based indices as in the reference: i1 = i + 1, half_lon(l) =
      // (l - 1) dlon + dlon / 2 (mesh_mod.F90:181-183)
      const int i1 = i + 1;
      const double x0 = R * ((double)(i1 - 1) * g.dlon + 0.5 * g.dlon) * cl;
      double uk[4], xk;
      uk[0] = u[i];
      xk = x0 + dt / 2 * uk[0];
#pragma unroll
      for (int st = 1; st < 4; st++) {
        const int l = (int)floor((double)i1 + (xk - x0) / dx);
        const double xl = R * ((double)(l - 1) * g.dlon + 0.5 * g.dlon) * cl;
        const double xr = R * ((double)(l + 1 - 1) * g.dlon + 0.5 * g.dlon) * cl;
        uk[st] = ((xr - xk) * u[wrap_any(l - 1, n)] + (xk - xl) * u[wrap_any(l, n)]) / dx;
        xk = st < 2 ? x0 + dt / 2 * uk[st] : x0 + dt * uk[st];
      }
      const double c = (uk[0] + 2 * uk[1] + 2 * uk[2] + uk[3]) / 6 * dt / dx;
      p.cflx[r * g.pitch + i] = c;
      if (fabs(c) > g.hw) bad |= F_CFLX;
      p.divx[r * g.pitch + i] = (u[i] - u[wrap(i - 1, n)]) * M(GMC_LE_LON, r) / M(GMC_AREA_CELL, r);
      p.divy[r * g.pitch + i] = (p.v[r * g.pitch + i] * M(GMC_LE_LAT, r) -
                                 p.v[(r - 1) * g.pitch + i] * M(GMC_LE_LAT, r - 1)) / M(GMC_AREA_CELL, r);
    }
  }
  if (r < g.nhalf) {
    const double dy = M(GMC_DE_LAT, r), y0 = R * M(GMC_HALF_LAT, r);
    for (int i = threadIdx.x; i < n; i += NA) {
      const int j1 = r + 1;
      double vk[4], yk;
      vk[0] = p.v[r * g.pitch + i];
      yk = y0 + dt / 2 * vk[0];
#pragma unroll
      for (int st = 1; st < 4; st++) {
        const int l = (int)floor((double)j1 + (yk - y0) / dy);
        if (l < 1 || l > g.nhalf - 1) {
          vk[st] = vk[0];
        } else {
          vk[st] = ((R * M(GMC_HALF_LAT, l) - yk) * p.v[(l - 1) * g.pitch + i] +
                    (yk - R * M(GMC_HALF_LAT, l - 1)) * p.v[l * g.pitch + i]) / dy;
        }
        yk = st < 2 ? y0 + dt / 2 * vk[st] : y0 + dt * vk[st];
      }
      const double c = (vk[0] + 2 * vk[1] + 2 * vk[2] + vk[3]) / 6 * dt / dy;
      p.cfly[r * g.pitch + i] = c;
      if (fabs(c) > 1) bad |= F_CFLY;
    }
  }
  if (r == 0 || r == g.nlat - 1) {
    const int h = r == 0 ? 0 : g.nhalf - 1;
    const double sgn = r == 0 ? 1.0 : -1.0;
    for (int i = threadIdx.x; i < n; i += NA) sv[i] = sgn * p.v[h * g.pitch + i];
    double pole = block_zsum(g, sv, &sv[4096]);
    pole = pole * M(GMC_LE_LAT, h) / n / M(GMC_AREA_CELL, r);
    for (int i = threadIdx.x; i < n; i += NA) p.divy[r * g.pitch + i] = pole;
  }
  if (bad) atomicOr(p.flags, bad);
}

// ---------------------------------------------------------------------------
struct InnerPtrs {
  const double *q;       // level old, tracer 0
  double *qx, *qy;       // tracer 0
  const double *u, *v, *cflx, *cfly, *divx, *divy;
  int64_t tstride;       // doubles between tracers
};

// ffsl_calc_tracer_hflx_cell, inner part (ffsl_mod.F90:350-407) for full
// row r of tracer blockIdx.y
__global__ void __launch_bounds__(NA) k_adv_inner(AdvGrid g, Mesh M, int lim, InnerPtrs p) {
  extern __shared__ double lds[];
  const int n = g.nlon, r = blockIdx.x;
  const int64_t toff = blockIdx.y * p.tstride;
  const double *q = p.q + toff;
  double *srow = lds, *sml = lds + n, *sdm = lds + 2 * n, *s6 = lds + 3 * n, *sfx = lds + 4 * n;
  double *sb = lds + 5 * n;
  const double dt = g.dt;
  for (int i = threadIdx.x; i < n; i += NA) srow[i] = q[r * g.pitch + i];
  __syncthreads();
  if (r > 0 && r < g.nlat - 1) {
    ppm_row(g, lim, srow, sml, sdm, s6);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += NA)
      sfx[i] = xflux(g, p.cflx[r * g.pitch + i], p.u[r * g.pitch + i], i, srow, sml, sdm, s6);
    __syncthreads();
    const double lel = M(GMC_LE_LON, r), ac = M(GMC_AREA_CELL, r);
    const double ls = M(GMC_LE_LAT, r - 1), ln = M(GMC_LE_LAT, r);
    for (int i = threadIdx.x; i < n; i += NA) {
      const double fs = yflux(g, lim, p.cfly[(r - 1) * g.pitch + i], p.v[(r - 1) * g.pitch + i], r - 1, i, q);
      const double fn = yflux(g, lim, p.cfly[r * g.pitch + i], p.v[r * g.pitch + i], r, i, q);
      const double qc = srow[i];
      const int64_t k = r * g.pitch + i;
      p.qx[toff + k] = qc - 0.5 * ((sfx[i] - sfx[wrap(i - 1, n)]) * lel / ac - p.divx[k] * qc) * dt;
      p.qy[toff + k] = qc - 0.5 * ((fn * ln - fs * ls) / ac - p.divy[k] * qc) * dt;
    }
  } else {
    // pole rows: qx = q, qy from the zonal sum of the adjacent V row's flux
    const int h = r == 0 ? 0 : g.nhalf - 1;
    for (int i = threadIdx.x; i < n; i += NA)
      sml[i] = yflux(g, lim, p.cfly[h * g.pitch + i], p.v[h * g.pitch + i], h, i, q);
    double pole = block_zsum(g, sml, sb);
    pole = pole * M(GMC_LE_LAT, h) / n / M(GMC_AREA_CELL, r);
    for (int i = threadIdx.x; i < n; i += NA) {
      const double qc = srow[i];
      const int64_t k = r * g.pitch + i;
      p.qx[toff + k] = qc;
      p.qy[toff + k] = r == 0 ? qc - 0.5 * (pole - p.divy[k] * qc) * dt
                              : qc + 0.5 * (pole - p.divy[k] * qc) * dt;
    }
  }
}

struct OuterPtrs {
  const double *qx, *qy;   // tracer 0
  const double *q;         // level old, tracer 0
  double *qn;              // level new, tracer 0
  const double *mfx, *mfy, *cflx, *cfly, *old_m, *m, *wgt;
  int64_t tstride;
  int *flags;
  unsigned long long *fixes;
  int damp_passes;
};

// outer fluxes + adv_run's update of one row (ffsl_mod.F90:410,
// adv_mod.F90:183-301) for full row r of tracer blockIdx.y
__global__ void __launch_bounds__(NA) k_adv_outer(AdvGrid g, Mesh M, int lim, OuterPtrs p) {
  extern __shared__ double lds[];
  const int n = g.nlon, r = blockIdx.x;
  const int64_t toff = blockIdx.y * p.tstride;
  const double *qx = p.qx + toff, *qy = p.qy + toff, *q = p.q + toff;
  double *qn = p.qn + toff;
  double *srow = lds, *sml = lds + n, *sdm = lds + 2 * n, *s6 = lds + 3 * n, *sfx = lds + 4 * n;
  double *sb = lds + 5 * n;  // [0] pole sum, [1] seam west, [2] seam east, [3] flag
  const double dt = g.dt;
  // the longitude halo of the new level still holds the values of its last
  // fill_halo: the row's own last and first values, read before overwriting
  if (threadIdx.x == 0) {
    sb[1] = qn[r * g.pitch + n - 1];
    sb[2] = qn[r * g.pitch];
  }
  if (r > 0 && r < g.nlat - 1) {
    for (int i = threadIdx.x; i < n; i += NA) srow[i] = qy[r * g.pitch + i];
    __syncthreads();
    ppm_row(g, lim, srow, sml, sdm, s6);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += NA)
      sfx[i] = xflux(g, p.cflx[r * g.pitch + i], p.mfx[r * g.pitch + i], i, srow, sml, sdm, s6);
    __syncthreads();
    const double lel = M(GMC_LE_LON, r), ac = M(GMC_AREA_CELL, r);
    const double ls = M(GMC_LE_LAT, r - 1), ln = M(GMC_LE_LAT, r);
    for (int i = threadIdx.x; i < n; i += NA) {
      const double fs = yflux(g, lim, p.cfly[(r - 1) * g.pitch + i], p.mfy[(r - 1) * g.pitch + i], r - 1, i, qx);
      const double fn = yflux(g, lim, p.cfly[r * g.pitch + i], p.mfy[r * g.pitch + i], r, i, qx);
      const int64_t k = r * g.pitch + i;
      const double qv = p.old_m[k] * q[k] -
                        ((sfx[i] - sfx[wrap(i - 1, n)]) * lel + (fn * ln - fs * ls)) / ac * dt;
      sml[i] = qv / p.m[k];  // the x-flux coefficients are no longer needed
    }
  } else {
    const int h = r == 0 ? 0 : g.nhalf - 1;
    for (int i = threadIdx.x; i < n; i += NA)
      sml[i] = yflux(g, lim, p.cfly[h * g.pitch + i], p.mfy[h * g.pitch + i], h, i, qx);
    double pole = block_zsum(g, sml, sb);
    pole = pole * M(GMC_LE_LAT, h) / n / M(GMC_AREA_CELL, r) * dt;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += NA) {
      const int64_t k = r * g.pitch + i;
      const double qv = r == 0 ? p.old_m[k] * q[k] - pole : p.old_m[k] * q[k] + pole;
      sml[i] = qv / p.m[k];
    }
  }
  // positivity fixer (adv_mod.F90:236-258): sequential along the row, run
  // by one lane only when the row holds a negative value
  double *sq = sml;
  const double *mrow = p.m + r * g.pitch;
  __syncthreads();
  if (threadIdx.x == 0) sb[3] = 0.0;
  __syncthreads();
  int neg = 0;
  for (int i = threadIdx.x; i < n; i += NA) neg |= sq[i] < 0;
  if (neg) sb[3] = 1.0;
  __syncthreads();
  if (sb[3] != 0.0 && threadIdx.x == 0) {
    double west = sb[1], east = sb[2];
    unsigned long long nfix = 0;
    int bad = 0;
    for (int i = 0; i < n; i++) {
      if (!(sq[i] < 0)) continue;
      const int im = i == 0 ? n - 1 : i - 1, ip = i == n - 1 ? 0 : i + 1;
      const double qm0 = sq[i] * mrow[i];
      const double qm1 = (i == 0 ? west : sq[im]) * mrow[im];
      const double qm2 = (i == n - 1 ? east : sq[ip]) * mrow[ip];
      const double qm0_half = 0.5 * qm0;
      if (qm1 >= qm0_half && qm2 >= qm0_half) {
        const double a = (qm1 - qm0_half) / mrow[im], b = (qm2 - qm0_half) / mrow[ip];
        if (i == 0) west = a; else sq[im] = a;
        if (i == n - 1) east = b; else sq[ip] = b;
      } else if (qm1 > qm0) {
        const double a = (qm1 - qm0) / mrow[im];
        if (i == 0) west = a; else sq[im] = a;
      } else if (qm2 > qm0) {
        const double b = (qm2 - qm0) / mrow[ip];
        if (i == n - 1) east = b; else sq[ip] = b;
      } else {
        bad = 1;
      }
      sq[i] = 0;
      nfix++;
    }
    atomicAdd(p.fixes, nfix);
    if (bad) atomicOr(p.flags, F_NEG);
  }
  __syncthreads();
  // lon_damp_on_cell, order 2 (lon_damp_mod.F90:158-241): Jacobi sweeps of
  // the row with periodic neighbours
  const double w = (r > 0 && r < g.nlat - 1) ? p.wgt[r] : 0.0;
  if (w != 0.0) {
    double *a = sq, *b = sdm;
    for (int pass = 0; pass < p.damp_passes; pass++) {
      for (int i = threadIdx.x; i < n; i += NA)
        b[i] = w * (a[i] + 0.25 * (a[wrap(i - 1, n)] - 2.0 * a[i] + a[wrap(i + 1, n)])) + (1 - w) * a[i];
      __syncthreads();
      double *t = a; a = b; b = t;
    }
    sq = a;
  }
  for (int i = threadIdx.x; i < n; i += NA) qn[r * g.pitch + i] = sq[i];
}

// row sums of q * area_cell of every tracer (the driver's diagnose)
__global__ void __launch_bounds__(NA) k_adv_mass(AdvGrid g, Mesh M, const double *q, int64_t tstride,
                                                 double *rows) {
  __shared__ double sw[NA / kWave];
  const int r = blockIdx.x;
  const double *qr = q + blockIdx.y * tstride + r * g.pitch;
  double acc = 0.0;
  for (int i = threadIdx.x; i < g.nlon; i += NA) acc = acc + qr[i] * M(GMC_AREA_CELL, r);
  for (int w = 1; w < kWave; w <<= 1) acc = acc + __shfl_xor(acc, w, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) sw[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < NA / kWave; k++) s = s + sw[k];
    rows[blockIdx.y * g.nlat + r] = s;
  }
}

}  // namespace
}  // namespace gmc

using namespace gmc;

struct gmc_adv {
  AdvGrid g{};
  gmc_adv_config cfg{};
  double *d_mesh = nullptr, *d_wgt = nullptr;
  double *slab = nullptr;
  size_t slab_bytes = 0;
  // block%state(itime) fields, wind basis, batch arrays
  double *su = nullptr, *sv = nullptr, *smfx = nullptr, *smfy = nullptr, *sm = nullptr;
  double *ub = nullptr, *vb = nullptr;
  double *u = nullptr, *v = nullptr, *mfx = nullptr, *mfy = nullptr, *old_m = nullptr;
  double *cflx = nullptr, *cfly = nullptr, *divx = nullptr, *divy = nullptr;
  double *q = nullptr, *qx = nullptr, *qy = nullptr;  // q: [2][ntracers][nlat][pitch]
  double *d_rows = nullptr;
  int *d_flags = nullptr;
  unsigned long long *d_fixes = nullptr;
  int *h_flags = nullptr;  // pinned
  hipStream_t stream = nullptr;
  int uv_step = 0, mf_step = 0, old = 1, new_ = 2;
  bool have_basis = false;
  bool old_m_current = false;  // old_m already equals m (copy_old_m can be skipped)
  size_t lds_bytes = 0;
  hipEvent_t *prof_ev = nullptr;
  int prof_n = 0, prof_cap = 0;
  std::vector<int> prof_kind;
  std::string err;
};

namespace {

int aerr(gmc_adv *a, int code, const std::string &msg) {
  if (a) a->err = msg;
  return code;
}
int ahip(gmc_adv *a, hipError_t e, const char *what) {
  if (e == hipSuccess) return GMC_OK;
  return aerr(a, e == hipErrorOutOfMemory ? GMC_ENOMEM : GMC_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

double *tracer_ptr(gmc_adv *a, int level, int l) {
  return a->q + ((int64_t)(level - 1) * a->cfg.ntracers + l) * a->g.nlat * a->g.pitch;
}

void mark(gmc_adv *a, int kind) {
  if (a->prof_ev && a->prof_n < a->prof_cap) {
    (void)hipEventRecord(a->prof_ev[a->prof_n++], a->stream);
    a->prof_kind.push_back(kind);
  }
}

int copy_rows(gmc_adv *a, double *dst, const double *host, int rows, int64_t ld) {
  return ahip(a, hipMemcpy2DAsync(dst, a->g.pitch * sizeof(double), host, ld * sizeof(double),
                                  a->g.nlon * sizeof(double), rows, hipMemcpyHostToDevice, a->stream),
              "upload");
}

// adv_run: accumulate, maybe advance, copy_old_m (adv_mod.F90:152-312)
int run_once(gmc_adv *a, int64_t time_step, bool basis, double scale) {
  const AdvGrid &g = a->g;
  const Mesh M{a->d_mesh, g.nlat};
  const int ns = a->cfg.nstep;
  // the two counters advance in step (adv_batch_mod.F90:238-251, 389-401)
  const int mode = a->uv_step == 0 ? 0 : (a->uv_step == ns ? 1 : 2);
  AccumPtrs ap{a->u, a->v, a->mfx, a->mfy, a->su, a->sv, a->smfx, a->smfy,
               basis ? a->ub : nullptr, basis ? a->vb : nullptr};
  hipLaunchKernelGGL(k_adv_accum, dim3(g.nlat), dim3(NA), 0, a->stream, g, ap, mode, ns, scale);
  mark(a, KA_ACCUM);
  a->uv_step = mode == 0 ? 1 : (mode == 1 ? 0 : a->uv_step + 1);
  a->mf_step = a->uv_step;
  if (a->uv_step == 0) {
    CflPtrs cp{a->u, a->v, a->cflx, a->cfly, a->divx, a->divy, a->d_flags};
    hipLaunchKernelGGL(k_adv_cfl, dim3(g.nlat), dim3(NA), 0, a->stream, g, M, cp);
    mark(a, KA_CFL);
  }
  if (time_step > 0 && time_step % ns == 0) {
    const int64_t ts = (int64_t)g.nlat * g.pitch;
    const dim3 grid(g.nlat, a->cfg.ntracers);
    InnerPtrs ip{tracer_ptr(a, a->old, 0), a->qx, a->qy, a->u, a->v, a->cflx, a->cfly, a->divx, a->divy, ts};
    hipLaunchKernelGGL(k_adv_inner, grid, dim3(NA), a->lds_bytes, a->stream, g, M, a->cfg.limiter, ip);
    mark(a, KA_INNER);
    OuterPtrs op{a->qx, a->qy, tracer_ptr(a, a->old, 0), tracer_ptr(a, a->new_, 0), a->mfx, a->mfy,
                 a->cflx, a->cfly, a->old_m, a->sm, a->d_wgt, ts, a->d_flags, a->d_fixes,
                 a->cfg.lon_damp_passes};
    hipLaunchKernelGGL(k_adv_outer, grid, dim3(NA), a->lds_bytes, a->stream, g, M, a->cfg.limiter, op);
    mark(a, KA_OUTER);
    std::swap(a->old, a->new_);
  }
  int rc = ahip(a, hipGetLastError(), "advection launch");
  if (!rc && !a->old_m_current)
    rc = ahip(a, hipMemcpyAsync(a->old_m, a->sm, (size_t)g.nlat * g.pitch * sizeof(double),
                                hipMemcpyDeviceToDevice, a->stream), "copy_old_m");
  if (!rc) a->old_m_current = true;
  return rc;
}

int check_flags(gmc_adv *a) {
  int rc = ahip(a, hipMemcpyAsync(a->h_flags, a->d_flags, sizeof(int), hipMemcpyDeviceToHost, a->stream), "flags");
  if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
  if (rc) return rc;
  const int f = *a->h_flags;
  if (f & F_CFLX) return aerr(a, GMC_ENUMERIC, "cflx exceeds lon_halo_width");
  if (f & F_CFLY) return aerr(a, GMC_ENUMERIC, "cfly exceeds 1");
  if (f & F_NEG) return aerr(a, GMC_ENUMERIC, "negative tracer the fixer cannot fill");
  return GMC_OK;
}

}  // namespace

extern "C" {

int gmc_adv_create(const gmc_mesh *mesh, const gmc_adv_config *cfg, gmc_adv **out) {
  if (!out) return GMC_EINVAL;
  *out = nullptr;
  if (!mesh || !cfg || !mesh->metrics || !cfg->lon_damp_wgt) return GMC_EINVAL;
  const int nlon = mesh->nlon, nlat = mesh->nlat;
  if (nlon < 8 || nlat < 5 || (nlon & 1) || nlon > 4096 || cfg->ntracers < 1 || cfg->nstep < 1 ||
      cfg->lon_halo_width < 1 || cfg->lon_halo_width > nlon / 2 - 4 || cfg->limiter < 0 ||
      cfg->limiter > 2 || cfg->lon_damp_passes < 0 || !(cfg->dt_adv > 0) || !(cfg->radius > 0))
    return GMC_EINVAL;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return GMC_EDEVICE;
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0) return GMC_EDEVICE;
  gmc_adv *a = new gmc_adv();
  a->cfg = *cfg;
  a->cfg.lon_damp_wgt = nullptr;
  AdvGrid &g = a->g;
  g.nlon = nlon;
  g.nlat = nlat;
  g.nhalf = nlat - 1;
  g.half = nlon / 2;
  g.pitch = (nlon + 31) / 32 * 32;
  g.hw = cfg->lon_halo_width;
  g.dlon = mesh->dlon;
  g.radius = cfg->radius;
  g.dt = cfg->dt_adv;
  a->lds_bytes = sizeof(double) * (5 * (size_t)nlon + 8);
  const size_t fld = (size_t)nlat * g.pitch;
  const int ntr = cfg->ntracers;
  // 16 single fields + q (2 levels) + qx, qy per tracer
  const size_t nd = fld * (16 + 4 * (size_t)ntr) + GMC_MESH_NFIELDS * (size_t)nlat + nlat +
                    (size_t)ntr * nlat;
  a->slab_bytes = nd * sizeof(double) + 256;
  int rc = ahip(a, hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking), "stream");
  if (!rc) rc = ahip(a, hipMalloc((void **)&a->slab, a->slab_bytes), "alloc");
  if (!rc) rc = ahip(a, hipMemset(a->slab, 0, a->slab_bytes), "memset");
  if (!rc) rc = ahip(a, hipHostMalloc((void **)&a->h_flags, sizeof(int)), "pinned");
  if (rc) {
    gmc_adv_destroy(a);
    return rc;
  }
  double *p = a->slab;
  double **f[] = {&a->su, &a->sv, &a->smfx, &a->smfy, &a->sm, &a->ub, &a->vb, &a->u, &a->v,
                  &a->mfx, &a->mfy, &a->old_m, &a->cflx, &a->cfly, &a->divx, &a->divy};
  for (auto pp : f) { *pp = p; p += fld; }
  a->q = p; p += 2 * fld * ntr;
  a->qx = p; p += fld * ntr;
  a->qy = p; p += fld * ntr;
  a->d_mesh = p; p += GMC_MESH_NFIELDS * (size_t)nlat;
  a->d_wgt = p; p += nlat;
  a->d_rows = p; p += (size_t)ntr * nlat;
  a->d_flags = (int *)p;
  a->d_fixes = (unsigned long long *)(p + 2);
  rc = ahip(a, hipMemcpy(a->d_mesh, mesh->metrics, sizeof(double) * GMC_MESH_NFIELDS * nlat, hipMemcpyHostToDevice), "mesh");
  if (!rc) rc = ahip(a, hipMemcpy(a->d_wgt, cfg->lon_damp_wgt, sizeof(double) * nlat, hipMemcpyHostToDevice), "weights");
  if (!rc) rc = ahip(a, hipFuncSetAttribute((const void *)k_adv_inner, hipFuncAttributeMaxDynamicSharedMemorySize, (int)a->lds_bytes), "lds");
  if (!rc) rc = ahip(a, hipFuncSetAttribute((const void *)k_adv_outer, hipFuncAttributeMaxDynamicSharedMemorySize, (int)a->lds_bytes), "lds");
  if (rc) {
    gmc_adv_destroy(a);
    return rc;
  }
  a->old = 1;
  a->new_ = 2;
  *out = a;
  return GMC_OK;
}

int gmc_adv_destroy(gmc_adv *a) {
  if (!a) return GMC_EINVAL;
  if (a->stream) (void)hipStreamSynchronize(a->stream);
  if (a->slab) (void)hipFree(a->slab);
  if (a->h_flags) (void)hipHostFree(a->h_flags);
  if (a->stream) (void)hipStreamDestroy(a->stream);
  delete a;
  return GMC_OK;
}

const char *gmc_adv_last_error(const gmc_adv *a) { return a ? a->err.c_str() : "null object"; }

int gmc_adv_upload_tracer(gmc_adv *a, int32_t level, int32_t tracer, const double *host, int64_t ld) {
  if (!a) return GMC_EINVAL;
  if (level < 1 || level > 2 || tracer < 0 || tracer >= a->cfg.ntracers || !host || ld < a->g.nlon)
    return aerr(a, GMC_EINVAL, "bad tracer upload arguments");
  int rc = copy_rows(a, tracer_ptr(a, level, tracer), host, a->g.nlat, ld);
  if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
  return rc;
}

int gmc_adv_download_tracer(gmc_adv *a, int32_t level, int32_t tracer, double *host, int64_t ld) {
  if (!a) return GMC_EINVAL;
  if (level < 1 || level > 2 || tracer < 0 || tracer >= a->cfg.ntracers || !host || ld < a->g.nlon)
    return aerr(a, GMC_EINVAL, "bad tracer download arguments");
  int rc = ahip(a, hipMemcpy2DAsync(host, ld * sizeof(double), tracer_ptr(a, level, tracer),
                                    a->g.pitch * sizeof(double), a->g.nlon * sizeof(double), a->g.nlat,
                                    hipMemcpyDeviceToHost, a->stream), "download");
  if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
  return rc;
}

int gmc_adv_set_wind(gmc_adv *a, const double *u, const double *v, const double *mfx, const double *mfy,
                     const double *m, int64_t ld) {
  if (!a) return GMC_EINVAL;
  if (ld < a->g.nlon) return aerr(a, GMC_EINVAL, "ld < nlon");
  int rc = GMC_OK;
  if (!rc && u) rc = copy_rows(a, a->su, u, a->g.nlat, ld);
  if (!rc && v) rc = copy_rows(a, a->sv, v, a->g.nhalf, ld);
  if (!rc && mfx) rc = copy_rows(a, a->smfx, mfx, a->g.nlat, ld);
  if (!rc && mfy) rc = copy_rows(a, a->smfy, mfy, a->g.nhalf, ld);
  if (!rc && m) {
    rc = copy_rows(a, a->sm, m, a->g.nlat, ld);
    a->old_m_current = false;
  }
  if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
  return rc;
}

int gmc_adv_run(gmc_adv *a, int64_t time_step) {
  if (!a) return GMC_EINVAL;
  if (time_step < 0) return aerr(a, GMC_EINVAL, "time_step < 0");
  int rc = run_once(a, time_step, false, 1.0);
  if (!rc) rc = check_flags(a);
  return rc;
}

int gmc_adv_set_wind_basis(gmc_adv *a, const double *ub, const double *vb, int64_t ld) {
  if (!a) return GMC_EINVAL;
  if (!ub || !vb || ld < a->g.nlon) return aerr(a, GMC_EINVAL, "bad wind basis");
  const AdvGrid &g = a->g;
  // set_uv writes u on the rows between the poles only
  int rc = ahip(a, hipMemsetAsync(a->ub, 0, (size_t)g.nlat * g.pitch * sizeof(double), a->stream), "basis");
  if (!rc) rc = copy_rows(a, a->ub + g.pitch, ub + ld, g.nlat - 2, ld);
  if (!rc) rc = copy_rows(a, a->vb, vb, g.nhalf, ld);
  if (!rc) {
    std::vector<double> one((size_t)g.nlat * g.pitch, 1.0);
    rc = ahip(a, hipMemcpyAsync(a->sm, one.data(), one.size() * sizeof(double), hipMemcpyHostToDevice, a->stream), "m");
    a->old_m_current = false;
    if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
  }
  if (!rc) a->have_basis = true;
  return rc;
}

int gmc_adv_drive(gmc_adv *a, int64_t time_step0, const double *scale, int32_t nsteps) {
  if (!a) return GMC_EINVAL;
  if (!a->have_basis) return aerr(a, GMC_ESTATE, "gmc_adv_set_wind_basis must come first");
  if (!scale || nsteps < 0 || time_step0 < 0) return aerr(a, GMC_EINVAL, "bad drive arguments");
  int rc = GMC_OK;
  for (int s = 0; s < nsteps && !rc; s++) rc = run_once(a, time_step0 + s, true, scale[s]);
  if (!rc) rc = check_flags(a);
  return rc;
}

int gmc_adv_old_level(const gmc_adv *a, int32_t *old_level) {
  if (!a || !old_level) return GMC_EINVAL;
  *old_level = a->old;
  return GMC_OK;
}

int gmc_adv_mass(gmc_adv *a, int32_t level, double *out) {
  if (!a) return GMC_EINVAL;
  if (level < 1 || level > 2 || !out) return aerr(a, GMC_EINVAL, "bad mass arguments");
  const AdvGrid &g = a->g;
  const int ntr = a->cfg.ntracers;
  hipLaunchKernelGGL(k_adv_mass, dim3(g.nlat, ntr), dim3(NA), 0, a->stream, g, Mesh{a->d_mesh, g.nlat},
                     tracer_ptr(a, level, 0), (int64_t)g.nlat * g.pitch, a->d_rows);
  std::vector<double> rows((size_t)ntr * g.nlat);
  int rc = ahip(a, hipGetLastError(), "mass launch");
  if (!rc) rc = ahip(a, hipMemcpyAsync(rows.data(), a->d_rows, rows.size() * sizeof(double), hipMemcpyDeviceToHost, a->stream), "mass copy");
  if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
  if (rc) return rc;
  for (int l = 0; l < ntr; l++) {
    double s = 0.0;
    for (int r = 0; r < g.nlat; r++) s += rows[(size_t)l * g.nlat + r];
    out[l] = s;
  }
  return GMC_OK;
}

int gmc_adv_fix_count(gmc_adv *a, int64_t *count) {
  if (!a || !count) return GMC_EINVAL;
  unsigned long long c = 0;
  int rc = ahip(a, hipMemcpyAsync(&c, a->d_fixes, sizeof(c), hipMemcpyDeviceToHost, a->stream), "fix count");
  if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
  *count = (int64_t)c;
  return rc;
}

int gmc_adv_synchronize(gmc_adv *a) {
  if (!a) return GMC_EINVAL;
  return ahip(a, hipStreamSynchronize(a->stream), "sync");
}

int gmc_adv_stream(gmc_adv *a, void **stream) {
  if (!a || !stream) return GMC_EINVAL;
  *stream = (void *)a->stream;
  return GMC_OK;
}

int gmc_adv_kernel_times(gmc_adv *a, int64_t time_step0, const double *scale, int32_t nsteps,
                         double *out_ms, int32_t cap, int32_t *n) {
  if (!a || !out_ms || !n || cap < 1) return GMC_EINVAL;
  if (!a->have_basis) return aerr(a, GMC_ESTATE, "gmc_adv_set_wind_basis must come first");
  if (!scale || nsteps < 1) return aerr(a, GMC_EINVAL, "bad timing arguments");
  std::vector<hipEvent_t> ev((size_t)nsteps * 5 + 1);
  int rc = GMC_OK;
  for (auto &e : ev)
    if (!rc) rc = ahip(a, hipEventCreate(&e), "event");
  if (rc) return rc;
  a->prof_ev = ev.data();
  a->prof_cap = (int)ev.size();
  a->prof_n = 0;
  a->prof_kind.clear();
  mark(a, -1);
  for (int s = 0; s < nsteps && !rc; s++) rc = run_once(a, time_step0 + s, true, scale[s]);
  a->prof_ev = nullptr;
  if (!rc) rc = check_flags(a);
  double acc[KA_NCLASS] = {};
  int cnt[KA_NCLASS] = {};
  for (int k = 1; !rc && k < a->prof_n; k++) {
    float ms = 0.f;
    rc = ahip(a, hipEventElapsedTime(&ms, ev[k - 1], ev[k]), "elapsed");
    acc[a->prof_kind[k]] += ms;
    cnt[a->prof_kind[k]]++;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  *n = std::min(cap, (int32_t)KA_NCLASS);
  for (int k = 0; k < *n; k++) out_ms[k] = cnt[k] ? acc[k] / cnt[k] : 0.0;
  return GMC_OK;
}

const char *gmc_adv_kernel_name(int32_t k) { return k >= 0 && k < KA_NCLASS ? kAdvNames[k] : ""; }

}  // extern "C"
