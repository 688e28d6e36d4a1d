// gmc_hflx.hpp -- the horizontal FFSL face fluxes (hflx_ppm_cell,
// hflx_van_leer_cell, ffsl_mod.F90:556-616, :778-881) of one field on one
// level, read from global memory: shared by the tracer advection kernels and
// the potential-temperature fluxes of the column operators.
//
// A field is a dense row-major [rows][pitch] array of doubles; longitude
// neighbours wrap; rows beyond a pole are the mirrored rows with a
// half-circle shift (the reference's fill_halo on one rank).  Every formula
// keeps the reference's operation order; the build uses -ffp-contract=off.
#pragma once

#include "gmc_internal.hpp"
#include "gmc_ppm.hpp"

namespace gmc {

struct AdvGrid {
  int nlon, nlat, nhalf, half;
  int64_t pitch;
  int hw;        // lon_halo_width: bound on |cflx|
  double dlon, radius, dt;
};

// Element i of a row whose base pointer is uniform over the wavefront, with
// a 32-bit byte offset: the load or store takes the scalar base and a vector
// offset (no 64-bit address arithmetic per lane).  i * 8 < 2^32 for every
// row of a field.
__device__ __forceinline__ double ldu(const double *base, unsigned i) {
  return *(const double *)((const char *)base + i * (unsigned)sizeof(double));
}
__device__ __forceinline__ void stu(double *base, unsigned i, double v) {
  *(double *)((char *)base + i * (unsigned)sizeof(double)) = v;
}

// The integral of the PPM parabola over [s1, s2] of the upwind cell, plus
// the whole cells crossed along x (hflx_ppm_cell, ffsl_mod.F90:831-877)
__device__ __forceinline__ double ppm_part_x(double s1, double s2, double ml, double dm, double m6,
                                             double sum) {
  const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
  return sum + ml * ds1 + 0.5 * dm * ds2 + m6 * (ds2 / 2.0 - div_rn(ds3, 3.0, 1.0 / 3.0));
}
__device__ __forceinline__ double ppm_part_y(double s1, double s2, double ml, double dm, double m6) {
  const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
  return ml * ds1 + 0.5 * dm * ds2 + m6 * (ds2 / 2.0 - div_rn(ds3, 3.0, 1.0 / 3.0));
}

// mx[k0] + mx[k0 + 1] + ... + mx[k0 + cnt - 1] (columns modulo n), added
// from 0 in ascending order; the LDS reads are issued eight at a time ahead
// of the dependent additions
__device__ __forceinline__ double run_sum(const double *mx, int k0, int cnt, int n) {
  double sum = 0;
  int j = 0;
  for (; j + 8 <= cnt; j += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = ldu(mx, wrap(k0 + j + u, n));
#pragma unroll
    for (int u = 0; u < 8; u++) sum = sum + v[u];
  }
  for (; j < cnt; j++) sum = sum + ldu(mx, wrap(k0 + j, n));
  return sum;
}

// y-flux through V point (h, i) of the field my (hflx_ppm_cell :855-877),
// with the PPM coefficients of the upwind full row (h if c > 0, else h + 1)
// from the five cells of column i around it.  h is uniform over the
// wavefront: the six rows h-2 .. h+3 are read through row pointers (rows
// beyond a pole mirrored, column half a circle away) and the upwind five
// selected per lane.
__device__ __forceinline__ double yflux(const AdvGrid &g, int lim, double c, double v, int h, int i,
                                        const double *my) {
  const unsigned ih = wrap(i + g.half, g.nlon), ic = i;
  double f[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int r = h - 2 + k;
    const bool mir = r < 0 || r > g.nlat - 1;
    const int rr = r < 0 ? -r : (r > g.nlat - 1 ? 2 * (g.nlat - 1) - r : r);
    f[k] = ldu(my + rr * g.pitch, mir ? ih : ic);
  }
  // both wind directions on one path (no divergence within a wavefront):
  // c > 0: v * part(1 - c, 1) / c from row h; c < 0: -v * part(0, -c) / c
  // from row h + 1
  const bool up = c > 0;
  const Ppm p = ppm(lim, up ? f[0] : f[1], up ? f[1] : f[2], up ? f[2] : f[3], up ? f[3] : f[4], up ? f[4] : f[5]);
  const double fy = (up ? v : -v) * ppm_part_y(up ? 1 - c : 0, up ? 1 : -c, p.fl, p.df, p.f6) / c;
  return c == 0 ? 0 : fy;
}

// x-flux through U point i of row `row` (hflx_ppm_cell, ffsl_mod.F90:831-853),
// read from global memory: the whole cells crossed and the PPM coefficients
// (ppm along x, :822-827) of the upwind cell.  |int(cflx)| is clamped to
// the halo bound the accumulation step has already checked, so every index
// stays within one wrap of the row.
__device__ __forceinline__ double xflux_g(const AdvGrid &g, int lim, double c, double u, int i,
                                          const double *row) {
  const int n = g.nlon;
  // cells i-2 .. i+3: the PPM stencil of either upwind cell when |c| < 1,
  // read without waiting for c so that these loads and c's are in flight
  // together
  double w[6];
#pragma unroll
  for (int k = 0; k < 6; k++) w[k] = ldu(row, wrap(i - 2 + k, n));
  int ci = (int)c;
  const double cf = c - ci;
  ci = ci > g.hw ? g.hw : (ci < -g.hw ? -g.hw : ci);
  // both wind directions on one path (no divergence within a wavefront):
  // c > 0: upwind cell i - ci, whole cells i+1-ci .. i, u * part(1 - cf, 1) / c;
  // c < 0: upwind cell i - ci + 1, whole cells i+1 .. i-ci, -u * part(0, -cf) / c
  const bool up = c > 0;
  double f[5], sum = 0;
  if (ci == 0) {   // upwind cell i or i + 1, no whole cells
#pragma unroll
    for (int k = 0; k < 5; k++) f[k] = up ? w[k] : w[k + 1];
  } else {
    const int iu = wrap(up ? i - ci : i - ci + 1, n);
    sum = run_sum(row, up ? i + 1 - ci : i + 1, up ? ci : -ci, n);
#pragma unroll
    for (int k = 0; k < 5; k++) f[k] = ldu(row, wrap(iu - 2 + k, n));
  }
  const Ppm p = ppm(lim, f[0], f[1], f[2], f[3], f[4]);
  const double fx = (up ? u : -u) * ppm_part_x(up ? 1 - cf : 0, up ? 1 : -cf, p.fl, p.df, p.f6, sum) / c;
  return c == 0 ? 0 : fx;
}

// van Leer fluxes (hflx_van_leer_cell, ffsl_mod.F90:577-611) of the same
// stencils: along x the upwind cell's value plus half its limited slope
// times 1 - cf (c > 0) or minus it times 1 + cf (c < 0), scaled by cf, plus
// or minus the whole cells crossed, times u / c; along y the upwind cell's
// value plus half its limited slope times sign(c) - c, times v.
__device__ __forceinline__ double xflux_vl(const AdvGrid &g, int lim, double c, double u, int i,
                                           const double *row) {
  const int n = g.nlon;
  int ci = (int)c;
  const double cf = c - ci;
  ci = ci > g.hw ? g.hw : (ci < -g.hw ? -g.hw : ci);
  const bool up = c > 0;
  const int iu = wrap(up ? i - ci : i - ci + 1, n);
  const double sum = run_sum(row, up ? i + 1 - ci : i + 1, up ? ci : -ci, n);
  const double fu = ldu(row, iu);
  const double dm = slope(lim, ldu(row, wrap(iu - 1, n)), fu, ldu(row, wrap(iu + 1, n)));
  const double a = up ? fu + dm * 0.5 * (1 - cf) : fu - dm * 0.5 * (1 + cf);
  const double fx = u * (up ? cf * a + sum : cf * a - sum) / c;
  return c == 0 ? 0 : fx;
}
__device__ __forceinline__ double yflux_vl(const AdvGrid &g, int lim, double c, double v, int h, int i,
                                           const double *my) {
  const unsigned ih = wrap(i + g.half, g.nlon), ic = i;
  double f[4];   // full rows h-1 .. h+2, beyond a pole mirrored
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int r = h - 1 + k;
    const bool mir = r < 0 || r > g.nlat - 1;
    const int rr = r < 0 ? -r : (r > g.nlat - 1 ? 2 * (g.nlat - 1) - r : r);
    f[k] = ldu(my + rr * g.pitch, mir ? ih : ic);
  }
  const bool up = c > 0;   // upwind row h, else h + 1
  const double fu = up ? f[1] : f[2];
  const double dm = slope(lim, up ? f[0] : f[1], fu, up ? f[2] : f[3]);
  return v * (fu + dm * 0.5 * (copysign(1.0, c) - c));
}
template <bool VL>
__device__ __forceinline__ double xflux(const AdvGrid &g, int lim, double c, double u, int i, const double *row) {
  return VL ? xflux_vl(g, lim, c, u, i, row) : xflux_g(g, lim, c, u, i, row);
}
template <bool VL>
__device__ __forceinline__ double yflux_any(const AdvGrid &g, int lim, double c, double v, int h, int i,
                                            const double *my) {
  return VL ? yflux_vl(g, lim, c, v, h, i, my) : yflux(g, lim, c, v, h, i, my);
}

}  // namespace gmc
