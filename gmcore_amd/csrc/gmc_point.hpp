// gmc_point.hpp -- point formulas of the shallow-water operators.
//
// One inline function per reference expression, written in the source's
// left-to-right order so that the operator-level kernels and the fused stage
// kernels produce the same bits.  Row r is 0-based (reference j = r + 1).
#pragma once

#include "gmc_internal.hpp"

namespace gmc {

// calc_m, operators_mod.F90:530: m = (gz - gzs) / g
__device__ __forceinline__ double m_of(double gz, double gzs, double g) { return (gz - gzs) / g; }

// interp_cell_to_vtx, interp_mod.F90:455-459
template <class MT>
__device__ __forceinline__ double m_vtx_of(double m00, double m01, double m10, double m11,
                                           const MT &M, int r) {
  return M.div((m00 + m01) * M(GMC_AREA_SUBCELL2, r) + (m10 + m11) * M(GMC_AREA_SUBCELL1, r + 1),
               GMC_AREA_VTX, r);
}

// w1 * (a + b) + w2 * (c + d): mfx_lat / mfy_lon, operators_mod.F90:585-600
__device__ __forceinline__ double tangent_sum(double w1, double w2, double a, double b, double c,
                                              double d) {
  return w1 * (a + b) + w2 * (c + d);
}

// calc_ke, ke_scheme 1, operators_mod.F90:247-252.  uw/ue: u west/east of
// the cell; vs/vn: v on the south/north edge.
template <class MT>
__device__ __forceinline__ double ke_of(double uw, double ue, double vs, double vn, const MT &M,
                                        int r) {
  return M.div(M(GMC_AREA_LON_WEST, r) * (uw * uw) + M(GMC_AREA_LON_EAST, r) * (ue * ue) +
                   M(GMC_AREA_LAT_NORTH, r - 1) * (vs * vs) + M(GMC_AREA_LAT_SOUTH, r) * (vn * vn),
               GMC_AREA_CELL, r);
}

// calc_div, operators_mod.F90:370-375
template <class MT>
__device__ __forceinline__ double div_of(double ue, double uw, double vn, double vs, const MT &M,
                                         int r) {
  return M.div((ue * M(GMC_LE_LON, r) - uw * M(GMC_LE_LON, r)) +
                   (vn * M(GMC_LE_LAT, r) - vs * M(GMC_LE_LAT, r - 1)),
               GMC_AREA_CELL, r);
}

// calc_vor, operators_mod.F90:629-634 (vertex row r)
template <class MT>
__device__ __forceinline__ double vor_of(double us, double un, double ve, double vw, const MT &M,
                                         int r) {
  return M.div(us * M(GMC_DE_LON, r) - un * M(GMC_DE_LON, r + 1) + ve * M(GMC_DE_LAT, r) -
                   vw * M(GMC_DE_LAT, r),
               GMC_AREA_VTX, r);
}

// calc_pv, operators_mod.F90:697
template <class MT>
__device__ __forceinline__ double pv_of(double vor, const MT &M, int r, double m_vtx) {
  return (vor + M(GMC_HALF_F, r)) / m_vtx;
}

// interp_pv_upwind, operators_mod.F90:764-790.  t: wind along the stencil
// (v_lon for pv_lon, u_lat for pv_lat), n: the other component; lat_edge
// selects the operand order under the square root.  f1..f4: pv along the
// stencil; the centred fallback averages f2 and f3.
__device__ __forceinline__ double pv_upwind_edge(int order, double wgt, double t, double n,
                                                 int lat_edge, double f1, double f2, double f3,
                                                 double f4) {
  const double b = fabs(t) / (sqrt(lat_edge ? t * t + n * n : n * n + t * t) + kEps);
  const double dir = copysign(1.0, t);
  const double up = order == 1 ? upwind1(dir, wgt, f2, f3) : upwind3(dir, wgt, f1, f2, f3, f4);
  return b * up + (1 - b) * 0.5 * (f2 + f3);
}

// tvd (tvd_mod.F90:42-61) with the flux limiter `lim` (GMC_TVD_*: none,
// van_leer, mc; tvd_mod.F90:63-100): the pv of interp_pv_tvd
// (operators_mod.F90:797-832) at an edge with Courant number cfl, from pv
// fm1, f, fp1, fp2 along the stencil.  eps is the default-real literal
// 1.0e-6 widened to r8, as in the reference.
__device__ __forceinline__ double tvd_edge(int lim, double cfl, double fm1, double f, double fp1, double fp2) {
  const double eps = (double)1.0e-6f;
  const double r = cfl > 0 ? (f - fm1 + eps) / (fp1 - f + eps) : (fp2 - fp1 + eps) / (fp1 - f + eps);
  double C;
  if (lim == GMC_TVD_NONE) C = 1;
  else if (lim == GMC_TVD_VAN_LEER) C = (r + fabs(r)) / (1 + fabs(r));
  else C = fmax(0.0, fmin(fmin(2.0 * r, 0.5 * (1 + r)), 2.0));
  return 0.5 * (fp1 + f - ((1 - C) * copysign(1.0, cfl) + cfl * C) * (fp1 - f));
}

// pv on an edge for pv_scheme `scheme`.  `order` is upwind_order_pv for the
// upwind scheme and the tvd limiter for tvd; dt_le = dt / le of the edge's
// row (tvd's Courant number t * dt / le is formed as (t * dt) / le).
__device__ __forceinline__ double pv_edge(int scheme, int order, double wgt, double t, double n, int lat_edge,
                                          double dt, double le, double f1, double f2, double f3, double f4) {
  if (scheme == GMC_PV_MIDPOINT) return 0.5 * (f3 + f2);
  if (scheme == GMC_PV_TVD) return tvd_edge(order, t * dt / le, f1, f2, f3, f4);
  return pv_upwind_edge(order, wgt, t, n, lat_edge, f1, f2, f3, f4);
}

// Derivatives of pv_upwind_edge with respect to t, n and the four pv values
// on the branch the trajectory took (the tangent-linear and adjoint of
// interp_pv_upwind; |t| and sign(t) differentiated piecewise): dt, dn, df[4].
struct UpwD {
  double dt, dn, df[4];
};
__device__ __forceinline__ UpwD upwind_deriv(int pv_scheme, int order, double wgt, double t, double n,
                                             int lat_edge, double f1, double f2, double f3, double f4) {
  UpwD d;
  if (pv_scheme == GMC_PV_MIDPOINT) {
    d.dt = d.dn = 0.0;
    d.df[0] = d.df[3] = 0.0;
    d.df[1] = d.df[2] = 0.5;
    return d;
  }
  const double S = sqrt(lat_edge ? t * t + n * n : n * n + t * t);
  const double den = S + kEps;
  const double b = fabs(t) / den;
  const double dir = copysign(1.0, t);
  const double w = wgt;
  double up, u1, u2, u3, u4;
  if (order == 1) {
    up = upwind1(dir, w, f2, f3);
    u1 = 0.0;
    u2 = 0.5 + 0.5 * w * dir;
    u3 = 0.5 - 0.5 * w * dir;
    u4 = 0.0;
  } else {
    up = upwind3(dir, w, f1, f2, f3, f4);
    const double c31 = 7.0 / 12.0, c32 = -1.0 / 12.0, c33 = 1.0 / 12.0;
    u1 = c32 - c33 * w * dir;
    u2 = c31 + 3 * c33 * w * dir;
    u3 = c31 - 3 * c33 * w * dir;
    u4 = c32 + c33 * w * dir;
  }
  const double c = 0.5 * (f2 + f3);
  // b = |t| / (S + eps): db/dt = sign(t) / den - |t| t / (S den^2), db/dn = -|t| n / (S den^2)
  const double q = S > 0 ? fabs(t) / (S * den * den) : 0.0;
  const double dbdt = copysign(1.0, t) / den - q * t, dbdn = -q * n;
  d.dt = dbdt * (up - c);
  d.dn = dbdn * (up - c);
  d.df[0] = b * u1;
  d.df[1] = b * u2 + (1 - b) * 0.5;
  d.df[2] = b * u3 + (1 - b) * 0.5;
  d.df[3] = b * u4;
  return d;
}

// calc_coriolis scheme 1, operators_mod.F90:866-877 (V point, row r)
template <class MT>
__device__ __forceinline__ double qhu_of(const MT &M, int r, double pl, double mfx_w, double pvl_w,
                                         double mfx_e, double pvl_e, double mfx_nw, double pvl_nw,
                                         double mfx_ne, double pvl_ne) {
  return (M(GMC_HALF_TANGENT_WGT1, r) * (mfx_w * (pl + pvl_w) + mfx_e * (pl + pvl_e)) +
          M(GMC_HALF_TANGENT_WGT2, r) * (mfx_nw * (pl + pvl_nw) + mfx_ne * (pl + pvl_ne))) *
         0.5;
}

// calc_coriolis scheme 1, operators_mod.F90:887-898 (U point, row r)
template <class MT>
__device__ __forceinline__ double qhv_of(const MT &M, int r, double pl, double mfy_sw, double pvt_sw,
                                         double mfy_se, double pvt_se, double mfy_nw, double pvt_nw,
                                         double mfy_ne, double pvt_ne) {
  return (M(GMC_FULL_TANGENT_WGT1, r) * (mfy_sw * (pl + pvt_sw) + mfy_se * (pl + pvt_se)) +
          M(GMC_FULL_TANGENT_WGT2, r) * (mfy_nw * (pl + pvt_nw) + mfy_ne * (pl + pvt_ne))) *
         0.5;
}

// calc_grad_mf, operators_mod.F90:956-970
template <class MT>
__device__ __forceinline__ double dmfdlon_of(double mfx_e, double mfx_w, const MT &M, int r) {
  return M.div((mfx_e - mfx_w) * M(GMC_LE_LON, r), GMC_AREA_CELL, r);
}
template <class MT>
__device__ __forceinline__ double dmfdlat_of(double mfy_n, double mfy_s, const MT &M, int r) {
  return M.div(mfy_n * M(GMC_LE_LAT, r) - mfy_s * M(GMC_LE_LAT, r - 1), GMC_AREA_CELL, r);
}

// pgf_swm_run, pgf_swm_mod.F90:33, :40
template <class MT>
__device__ __forceinline__ double pgf_lon_of(double gz_e, double gz_w, const MT &M, int r) {
  return M.div(gz_e - gz_w, GMC_DE_LON, r);
}
template <class MT>
__device__ __forceinline__ double pgf_lat_of(double gz_n, double gz_s, const MT &M, int r) {
  return M.div(gz_n - gz_s, GMC_DE_LAT, r);
}

// pole-row v smoothing, time_schemes_mod.F90:225-233 (wgt = 0.6)
__device__ __forceinline__ double pole_smooth(double v, double v_next) {
  const double wgt = 0.6;
  return (1 - wgt) * v + wgt * v_next;
}

// one output of the zonal filter, filter_mod.F90:41: taps in ascending
// order from x[0] (= column i - hn).  The products of eight taps are formed
// ahead of their additions, which stay in tap order, so the LDS reads of a
// wide polar filter overlap instead of each waiting for the last.
__device__ __forceinline__ double filter_point(const double *__restrict__ w, int n, const double *x) {
  double acc = 0.0;
  int k = 0;
  for (; k + 8 <= n; k += 8) {
    double pr[8];
#pragma unroll
    for (int u = 0; u < 8; u++) pr[u] = w[k + u] * x[k + u];
#pragma unroll
    for (int u = 0; u < 8; u++) acc = acc + pr[u];
  }
  if (k + 4 <= n) {
    double pr[4];
#pragma unroll
    for (int u = 0; u < 4; u++) pr[u] = w[k + u] * x[k + u];
#pragma unroll
    for (int u = 0; u < 4; u++) acc = acc + pr[u];
    k += 4;
  }
  for (; k < n; k++) acc = acc + w[k] * x[k];
  return acc;
}

// Sum of f(i), i = 0..n-1, over one wavefront in the library's fixed order:
// lane l adds f(l), f(l+64), ... in ascending order, then an xor butterfly
// combines the 64 partials (pairs (0,1), then (0,2), ...).  Every lane
// returns the same value.  Stands in for zonal_sum (parallel_mod.F90:741-764)
// whose order the reference leaves to the compiler and MPI.
template <class F>
__device__ __forceinline__ double wave_zonal_sum(int n, int lane, F f) {
  double acc = 0.0;
  for (int i = lane; i < n; i += kWave) acc = acc + f(i);
  for (int w = 1; w < kWave; w <<= 1) acc = acc + __shfl_xor(acc, w, kWave);
  return acc;
}

// wave_zonal_sum of x[0 .. n-1] read from memory, with each lane's loads
// issued eight at a time ahead of its additions; the additions, and so the
// result, are those of wave_zonal_sum.
__device__ __forceinline__ double wave_zonal_sum_batched(int n, int lane, const double *x) {
  double acc = 0.0;
  for (int i = lane; i < n; i += 8 * kWave) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = i + u * kWave < n ? x[i + u * kWave] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (i + u * kWave < n) acc = acc + v[u];
  }
  for (int w = 1; w < kWave; w <<= 1) acc = acc + __shfl_xor(acc, w, kWave);
  return acc;
}

}  // namespace gmc
