// gmc_baro.hip -- hydrostatic column operators of the baroclinic dynamics
// (include/gmcore_baro.h).
//
// Fields are [levels][rows][pitch] arrays in one device slab; V-point fields
// have nlat - 1 rows per level.  The level-local operators give a thread one
// column and a chunk of LC consecutive levels (grid x over the points of a
// level plane, grid y over the chunks), so the loads and stores of a level
// are coalesced along the latitude row and a launch has LC-times fewer
// threads than points but enough waves to fill the chip at 32 levels.  The
// geopotential's sums stage a block of columns in LDS; the two level
// recurrences (the surface-pressure tendency and the vertical mass flux)
// keep one thread per column.  The arithmetic is the reference's, operation
// by operation and in its order (IEEE division, no contraction); pow and log
// are the device library's, within a few ulp of the host's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gmc_internal.hpp"
#include "gmc_point.hpp"
#include "gmc_hflx.hpp"
#include "gmc_ppm.hpp"
#include "../../include/gmcore_baro.h"

namespace gmc {
namespace {

constexpr int NB = 256;   // columns per workgroup

struct BaroGrid {
  int nlon, nlat, nlev;
  int64_t pitch;
  int64_t sc, sv;         // level strides (doubles) of cell / U-point and V-point fields
};

// levels of a field: 0 one level, 1 full levels, 2 half levels; V-point rows
struct FieldShape {
  int lev;
  bool vrows;
};
constexpr FieldShape kShape[GMB_NFIELDS] = {
    {0, false}, {0, false}, {0, false},                  // phs, gzs, dphs
    {1, false}, {2, false}, {2, false}, {1, false},      // pt, ph_lev, ph_exn_lev, ph
    {1, false}, {2, false}, {1, false},                  // t, gz_lev, gz
    {1, false}, {1, false},                              // dmfdlon, dmfdlat
    {2, false}, {2, false}, {2, true},                   // we_lev, we_lev_lon, we_lev_lat
    {1, false}, {1, true}, {1, false}, {1, true},        // u_f, v_f, m_lon, m_lat
    {1, false}, {1, true}, {1, false}, {1, true},        // wedudlev, wedvdlev, pgf_lon, pgf_lat
    {2, false}, {2, false}, {1, false},                  // cflz, ptf_lev, dptfdlev
    {1, false}, {1, true},                               // u_lon, v_lat
    {1, false}, {1, true}, {1, false}, {1, true},        // adv_u, adv_v, adv_mfx, adv_mfy
    {1, false}, {1, true}, {1, false}, {1, false},       // adv_cflx, adv_cfly, adv_divx, adv_divy
    {1, false}, {1, true}, {1, false}, {1, false}};      // ptf_lon, ptf_lat, dptfdlon, dptfdlat

constexpr int LC = 4;     // levels per thread of the level-chunked kernels
#ifndef GMC_POW_LC
#define GMC_POW_LC 2
#endif
constexpr int LCP = GMC_POW_LC;   // levels per thread of calc_ph and calc_t (a pow per level)
enum { kFlagCflz = 1, kFlagCflx = 2 };   // bits of the error flag: Courant numbers out of range

// How the blocks of a launch over (plane blocks, levels or level chunks) are
// dealt to the XCDs, each with its own L2 (xcd_item over the linear block
// index): kByRows gives an XCD a contiguous run of a level's plane blocks
// (neighbouring rows of a level share its L2: horizontal stencils), kByColumns
// a run of column blocks with all their levels (neighbouring levels of a
// column share it: vertical stencils), kAsLaunched the hardware's
// round-robin (pointwise work).
enum Deal { kAsLaunched, kByRows, kByColumns };

// The point of this thread in a launch over `rows` rows of a level plane
// (grid x) and the levels or level chunks (grid y): o = r * pitch + i, lev
// the level or chunk; false in the pitch padding and past the last row.
template <Deal D = kAsLaunched>
__device__ __forceinline__ bool plane_point(const BaroGrid &g, int rows, int64_t &o, int &i, int &r, int &lev) {
  const int gx = gridDim.x, gy = gridDim.y;
  int bx = blockIdx.x;
  lev = blockIdx.y;
  if (D != kAsLaunched) {
    const int it = xcd_item(blockIdx.x + blockIdx.y * gx, gx * gy);
    bx = D == kByRows ? it % gx : it / gy;
    lev = D == kByRows ? it / gx : it % gy;
  }
  const uint32_t p = (uint32_t)bx * NB + threadIdx.x;
  r = (int)(p / (uint32_t)g.pitch);
  i = (int)(p - (uint32_t)r * (uint32_t)g.pitch);
  o = p;
  return r < rows && i < g.nlon;
}

// calc_ph (operators_mod.F90:136-168): the half-level pressures of the
// hybrid coordinate (hybrid_coord_mod.F90:155-162), their Exner powers, and
// the full levels' mean of the two faces.  The chunk of full levels k0 .. k1-1
// writes their top half levels; the last chunk also the bottom one.
__device__ __forceinline__ double half_level_p(const double *hyai, const double *hybi, double p0, double ptop,
                                               double ps, int k) {
  return hyai[k] * (p0 - ptop) + hybi[k] * (ps - ptop) + ptop;
}

__global__ void __launch_bounds__(NB) k_calc_ph(BaroGrid g, const double *__restrict__ hyai,
                                                const double *__restrict__ hybi, double p0, double ptop,
                                                double rd_o_cpd, const double *__restrict__ phs, double *ph_lev,
                                                double *ph_exn_lev, double *ph) {
  int64_t o;
  int i, r, ch;
  if (!plane_point(g, g.nlat, o, i, r, ch)) return;
  const int k0 = ch * LCP, k1 = min(k0 + LCP, g.nlev);
  const double ps = phs[o];
  double above = half_level_p(hyai, hybi, p0, ptop, ps, k0);
  for (int k = k0; k < k1; k++) {
    const double below = half_level_p(hyai, hybi, p0, ptop, ps, k + 1);
    ph_lev[k * g.sc + o] = above;
    ph_exn_lev[k * g.sc + o] = pow(above, rd_o_cpd);
    ph[k * g.sc + o] = 0.5 * (above + below);
    above = below;
  }
  if (k1 == g.nlev) {
    ph_lev[k1 * g.sc + o] = above;
    ph_exn_lev[k1 * g.sc + o] = pow(above, rd_o_cpd);
  }
}

// calc_t (operators_mod.F90:170-190) with temperature() (formula_mod.F90:24-31)
__global__ void __launch_bounds__(NB) k_calc_t(BaroGrid g, double p0, double rd_o_cpd, const double *__restrict__ pt,
                                               const double *__restrict__ ph, double *t) {
  int64_t o;
  int i, r, ch;
  if (!plane_point(g, g.nlat, o, i, r, ch)) return;
  const int k0 = ch * LCP;
  double a[LCP], b[LCP];
#pragma unroll
  for (int j = 0; j < LCP; j++) {   // every load of the chunk in flight at once
    const int k = min(k0 + j, g.nlev - 1);
    a[j] = pt[k * g.sc + o];
    b[j] = ph[k * g.sc + o];
  }
#pragma unroll
  for (int j = 0; j < LCP; j++)
    if (k0 + j < g.nlev) t[(k0 + j) * g.sc + o] = a[j] * pow(b[j] / p0, rd_o_cpd);
}

// calc_gz_lev (operators_mod.F90:433-469).  Each half level's sum runs from
// 0 over the levels below it in ascending order, as the reference's inner
// loop does (so no running suffix sum: that would change the bits).  A block
// holds cw columns (a power of two) x NB / cw level lanes: the layer terms
// Rd t log(ph_lev(l+1) / ph_lev(l)) of its columns go to LDS (term l of
// column c at lds[l * cw + c]), lane q sums the half levels q, q + NB / cw, ..
// out of LDS, GZB of them at a time in one ascending pass over the terms (a
// chain per half level, each in its own order), and the full-level means read
// the half levels back from a second LDS array (the bottom half level, which
// keeps its value, from HBM).  (Forming the terms in a launch of their own,
// a column and two levels per thread, and summing them in a second one was
// slower: 13.9 + 22.3 us against 29.9 us at 360 x 181 x 32.)
constexpr int GZB = 8;

__global__ void __launch_bounds__(NB) k_calc_gz_lev(BaroGrid g, int cw, double rd, const double *__restrict__ gzs,
                                                    const double *__restrict__ t,
                                                    const double *__restrict__ ph_lev, double *gz_lev, double *gz) {
  extern __shared__ double lds[];
  const int n = g.nlev, c = threadIdx.x & (cw - 1), q = threadIdx.x / cw, lanes = NB / cw;
  double *term = lds, *glev = lds + (size_t)n * cw;
  const uint32_t p = blockIdx.x * (uint32_t)cw + c;
  const int r = (int)(p / (uint32_t)g.pitch), i = (int)(p - (uint32_t)r * (uint32_t)g.pitch);
  const bool ok = r < g.nlat && i < g.nlon;
  const int64_t o = ok ? (int64_t)p : 0;   // a column outside the block's rows reads column 0, unused
  for (int l0 = q; l0 < n; l0 += GZB * lanes) {
    double tv[GZB], pa[GZB], pb[GZB];
#pragma unroll
    for (int j = 0; j < GZB; j++) {   // the batch's loads in flight at once
      const int l = min(l0 + j * lanes, n - 1);
      tv[j] = t[l * g.sc + o];
      pa[j] = ph_lev[l * g.sc + o];
      pb[j] = ph_lev[(l + 1) * g.sc + o];
    }
#pragma unroll
    for (int j = 0; j < GZB; j++) {
      const int l = l0 + j * lanes;
      if (l < n) term[l * cw + c] = ok ? rd * tv[j] * log(pb[j] / pa[j]) : 0.0;
    }
  }
  __syncthreads();
  const double s = gzs[o];
  // lane q's half levels k0 + j lanes; chain j starts at its own level.  With
  // cw = 64 the lane is uniform over the wavefront, so the chains that have
  // not started are skipped by a scalar branch rather than selected away.
  const auto sums = [&](int qq) {
    for (int k0 = qq; k0 < n; k0 += GZB * lanes) {
      double acc[GZB];
#pragma unroll
      for (int j = 0; j < GZB; j++) acc[j] = 0.0;
      for (int l = k0; l < n; l++) {
        const double v = term[l * cw + c];
#pragma unroll
        for (int j = 0; j < GZB; j++)
          if (l >= k0 + j * lanes) acc[j] = acc[j] + v;
      }
#pragma unroll
      for (int j = 0; j < GZB; j++) {
        const int k = k0 + j * lanes;
        if (k < n) {
          const double v = s + acc[j];
          glev[k * cw + c] = v;
          if (ok) gz_lev[k * g.sc + o] = v;
        }
      }
    }
  };
  if (cw == kWave)
    sums(__builtin_amdgcn_readfirstlane(q));
  else
    sums(q);
  __syncthreads();
  if (!ok) return;
  for (int k = q; k < n; k += lanes) {
    const double below = k + 1 < n ? glev[(k + 1) * cw + c] : gz_lev[n * g.sc + o];
    gz[k * g.sc + o] = 0.5 * (glev[k * cw + c] + below);
  }
}

// calc_dphsdt (operators_mod.F90:1120-1143)
__global__ void __launch_bounds__(NB) k_calc_dphsdt(BaroGrid g, const double *__restrict__ dmfdlon,
                                                    const double *__restrict__ dmfdlat, double *dphs) {
  const int i = blockIdx.x * NB + threadIdx.x, r = blockIdx.y;
  if (i >= g.nlon) return;
  const int64_t o = r * g.pitch + i;
  double d = 0.0;
  for (int k0 = 0; k0 < g.nlev; k0 += 8) {
    double a[8], b[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {   // the batch's loads in flight at once
      const int k = min(k0 + j, g.nlev - 1);
      a[j] = dmfdlon[k * g.sc + o];
      b[j] = dmfdlat[k * g.sc + o];
    }
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (k0 + j < g.nlev) d = d - a[j] - b[j];
  }
  dphs[o] = d;
}

// calc_we_lev (operators_mod.F90:192-221): the mass-flux divergence above
// each interior half level accumulated in level order (the prefix the
// reference re-adds for every level has the same bits), minus
// dphdt_lev = hybi dphs (hybrid_coord_mod.F90:164-171); 0 at the top and
// the bottom
__global__ void __launch_bounds__(NB) k_calc_we_lev(BaroGrid g, const double *__restrict__ hybi,
                                                    const double *__restrict__ dmfdlon,
                                                    const double *__restrict__ dmfdlat,
                                                    const double *__restrict__ dphs, double *we_lev) {
  const int i = blockIdx.x * NB + threadIdx.x, r = blockIdx.y;
  if (i >= g.nlon) return;
  const int64_t o = r * g.pitch + i;
  const double dp = dphs[o];
  double mf = 0.0;
  we_lev[o] = 0.0;
  for (int k0 = 1; k0 < g.nlev; k0 += 8) {
    double a[8], b[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {   // the batch's loads in flight at once
      const int k = min(k0 + j, g.nlev - 1);
      a[j] = dmfdlon[(k - 1) * g.sc + o];
      b[j] = dmfdlat[(k - 1) * g.sc + o];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int k = k0 + j;
      if (k < g.nlev) {
        mf = mf + a[j] + b[j];
        we_lev[k * g.sc + o] = -(hybi[k] * dp) - mf;
      }
    }
  }
  we_lev[g.nlev * g.sc + o] = 0.0;
}

// The U rows 1 .. nlat-2 and the V rows 0 .. nlat-2 of one launch: virtual
// row y < nlat - 2 is U row y + 1, the rest V row y - (nlat - 2).
__device__ __forceinline__ bool u_or_v_row(const BaroGrid &g, int y, int &r) {
  const bool u = y < g.nlat - 2;
  r = u ? y + 1 : y - (g.nlat - 2);
  return u;
}

// interp_lev_edge_to_lev_lon_edge / _lat_edge with no wind
// (interp_mod.F90:360-369, :425-434), the chunk's half levels.  The thread
// of cell (i, r), r = 0 .. nlat-2, forms the V point north of it and, off the
// south pole row, the U point east of it: both read the cell's values.
__global__ void __launch_bounds__(NB) k_we_edges(BaroGrid g, Mesh M, const double *__restrict__ we_lev,
                                                 double *we_lon, double *we_lat) {
  int64_t o;
  int i, r, ch;
  if (!plane_point<kByRows>(g, g.nlat - 1, o, i, r, ch)) return;
  const int k0 = ch * LC, k1 = min(k0 + LC, g.nlev + 1);
  const int64_t oe = o - i + (i + 1 == g.nlon ? 0 : i + 1);
  double w[LC], e[LC], nn[LC];
#pragma unroll
  for (int j = 0; j < LC; j++) {   // the chunk's loads in flight at once
    const int64_t k = min(k0 + j, g.nlev) * g.sc;
    w[j] = we_lev[k + o];
    e[j] = we_lev[k + oe];
    nn[j] = we_lev[k + o + g.pitch];
  }
  if (r >= 1) {
    const double aw = M(GMC_AREA_LON_WEST, r), ae = M(GMC_AREA_LON_EAST, r), a = M(GMC_AREA_LON, r);
#pragma unroll
    for (int j = 0; j < LC; j++)
      if (k0 + j < k1) we_lon[(k0 + j) * g.sc + o] = (aw * w[j] + ae * e[j]) / a;
  }
  const double an = M(GMC_AREA_LAT_NORTH, r), as = M(GMC_AREA_LAT_SOUTH, r), a = M(GMC_AREA_LAT, r);
#pragma unroll
  for (int j = 0; j < LC; j++)
    if (k0 + j < k1) we_lat[(k0 + j) * g.sv + o] = (an * nn[j] + as * w[j]) / a;
}

// calc_wedudlev_wedvdlev (operators_mod.F90:1145-1216) on full levels
// k0 .. k0+LC-1 (those below nlev) of one column of one point kind: x the
// wind (u_f or v_f), we its half-level mass flux, m its full-level mass; the
// top and bottom levels take their one face.  The chunk's loads (x on
// k0-1 .. k0+LC, we on k0 .. k0+LC, m, indexes clamped into the column) are
// issued before any arithmetic.
__device__ __forceinline__ void wedx_levels(int nlev, int k0, int64_t s, const double *__restrict__ x,
                                            const double *__restrict__ we, const double *__restrict__ m,
                                            double *out) {
  double xv[LC + 2], wv[LC + 1], mv[LC];
#pragma unroll
  for (int j = 0; j < LC + 2; j++) xv[j] = x[min(max(k0 - 1 + j, 0), nlev - 1) * s];
#pragma unroll
  for (int j = 0; j < LC + 1; j++) wv[j] = we[min(k0 + j, nlev) * s];
#pragma unroll
  for (int j = 0; j < LC; j++) mv[j] = m[min(k0 + j, nlev - 1) * s];
#pragma unroll
  for (int j = 0; j < LC; j++) {
    const int k = k0 + j;
    if (k >= nlev) break;
    const double xm = xv[j], xk = xv[j + 1], xp = xv[j + 2];
    double v;
    if (k == 0)
      v = (wv[j + 1] * (xp - xk)) / mv[j] / 2.0;
    else if (k == nlev - 1)
      v = (wv[j] * (xk - xm)) / mv[j] / 2.0;
    else
      v = (wv[j + 1] * (xp - xk) + wv[j] * (xk - xm)) / mv[j] / 2.0;
    out[k * s] = v;
  }
}

__global__ void __launch_bounds__(NB) k_wedudlev(BaroGrid g, const double *__restrict__ u,
                                                 const double *__restrict__ we_lon, const double *__restrict__ m_lon,
                                                 double *wedudlev, const double *__restrict__ v,
                                                 const double *__restrict__ we_lat,
                                                 const double *__restrict__ m_lat, double *wedvdlev) {
  int64_t o;
  int i, y, r, ch;
  if (!plane_point<kByColumns>(g, 2 * g.nlat - 3, o, i, y, ch)) return;
  const int k0 = ch * LC;
  const bool isu = u_or_v_row(g, y, r);
  const int64_t c = r * g.pitch + i;
  if (isu)
    wedx_levels(g.nlev, k0, g.sc, u + c, we_lon + c, m_lon + c, wedudlev + c);
  else
    wedx_levels(g.nlev, k0, g.sv, v + c, we_lat + c, m_lat + c, wedvdlev + c);
}

// pgf_lin97_run, hydrostatic branch (pgf_lin97_mod.F90:59-105): corners 1 .. 4
// of the quadrilateral between the point's two cells (a: west / south, b:
// east / north) and its two half levels
__device__ __forceinline__ double lin97(double exn_a_top, double exn_a_bot, double exn_b_top, double exn_b_bot,
                                        double gz_a_top, double gz_a_bot, double gz_b_top, double gz_b_bot,
                                        double de) {
  const double dph1 = exn_b_bot - exn_a_top;   // 2 - 4
  const double dph2 = exn_a_bot - exn_b_top;   // 1 - 3
  const double dgz1 = gz_a_bot - gz_b_top;     // 1 - 3
  const double dgz2 = gz_a_top - gz_b_bot;     // 4 - 2
  return -(dph1 * dgz1 + dph2 * dgz2) / de / (dph1 + dph2);
}

// The thread of cell (i, r), r = 0 .. nlat-2, forms the V point north of it
// and, off the south pole row, the U point east of it on full levels
// k0 .. k0+LC-1 (those below nlev): the lower half level of level k is the
// upper one of level k + 1, and the chunk's half levels k0 .. k0+LC of the
// cell and its two neighbours are loaded (clamped to the column) before any
// arithmetic.
__global__ void __launch_bounds__(NB) k_pgf_lin97(BaroGrid g, Mesh M, const double *__restrict__ exn,
                                                  const double *__restrict__ gz, double *pgf_lon, double *pgf_lat) {
  int64_t o;
  int i, r, ch;
  if (!plane_point<kByRows>(g, g.nlat - 1, o, i, r, ch)) return;
  const int k0 = ch * LC;
  const int64_t oe = o - i + (i + 1 == g.nlon ? 0 : i + 1), on = o + g.pitch;
  double ea[LC + 1], ga[LC + 1], ee[LC + 1], ge[LC + 1], en[LC + 1], gn[LC + 1];
#pragma unroll
  for (int j = 0; j < LC + 1; j++) {
    const int64_t k = min(k0 + j, g.nlev) * g.sc;
    ea[j] = exn[k + o];
    ga[j] = gz[k + o];
    ee[j] = exn[k + oe];
    ge[j] = gz[k + oe];
    en[j] = exn[k + on];
    gn[j] = gz[k + on];
  }
  if (r >= 1) {
    const double de = M(GMC_DE_LON, r);
#pragma unroll
    for (int j = 0; j < LC; j++)
      if (k0 + j < g.nlev)
        pgf_lon[(k0 + j) * g.sc + o] = lin97(ea[j], ea[j + 1], ee[j], ee[j + 1], ga[j], ga[j + 1], ge[j], ge[j + 1], de);
  }
  const double de = M(GMC_DE_LAT, r);
#pragma unroll
  for (int j = 0; j < LC; j++)
    if (k0 + j < g.nlev)
      pgf_lat[(k0 + j) * g.sv + o] = lin97(ea[j], ea[j + 1], en[j], en[j + 1], ga[j], ga[j + 1], gn[j], gn[j + 1], de);
}

// pt on level j = -2 .. nlev+1 of a column: the levels themselves, and two
// beyond the top and the bottom extrapolated linearly from the two nearest
// (operators_mod.F90:1085-1088)
__device__ __forceinline__ double ext_pt(const double *__restrict__ p, int64_t s, int n, int j) {
  if (j >= 0 && j < n) return p[j * s];
  if (j < 0) {
    const double e1 = 2 * p[0] - p[s];
    return j == -1 ? e1 : 2 * e1 - p[0];
  }
  const double en = 2 * p[(n - 1) * s] - p[(n - 2) * s];
  return j == n ? en : 2 * en - p[(n - 1) * s];
}

// vflx_ppm_cell (ffsl_mod.F90:883-941) through interior half level k of a
// column: the upwind cell's PPM partial plus the whole cells the Courant
// number c crosses, times the mass flux w.  bad is set where c's upwind cell
// falls outside the column (the indexes are then clamped so that every access
// stays in the column's levels -2 .. nlev+1).
__device__ __forceinline__ double pt_face_flux(int lim, const double *__restrict__ p, int64_t s, int n, int k,
                                               double c, double w, bool &bad) {
  if (c == 0) return 0.0;
  int ci = (int)c;
  const double cf = c - ci;
  ci = ci > n + 4 ? n + 4 : (ci < -n - 4 ? -n - 4 : ci);   // beyond the column either way: flagged below
  const bool down = c > 0;
  // c > 0: upwind cell k - ci - 1, whole cells k - ci .. k - 1;
  // c < 0: upwind cell k - ci, whole cells k .. k - ci - 1
  int ku = down ? k - ci - 1 : k - ci;
  int j0 = down ? k - ci : k, j1 = down ? k - 1 : k - ci - 1;
  if (ku < 0 || ku > n - 1 || j0 < -2 || j1 > n + 1) {
    bad = true;
    ku = ku < 0 ? 0 : (ku > n - 1 ? n - 1 : ku);
    j0 = j0 < -2 ? -2 : j0;
    j1 = j1 > n + 1 ? n + 1 : j1;
  }
  double sum = 0.0;
  for (int j = j0; j <= j1; j++) sum = sum + ext_pt(p, s, n, j);
  const Ppm q = ppm(lim, ext_pt(p, s, n, ku - 2), ext_pt(p, s, n, ku - 1), ext_pt(p, s, n, ku),
                    ext_pt(p, s, n, ku + 1), ext_pt(p, s, n, ku + 2));
  const double s1 = down ? 1 - cf : 0.0, s2 = down ? 1.0 : -cf;
  const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
  return (down ? w : -w) * (sum + q.fl * ds1 + 0.5 * q.df * ds2 + q.f6 * (ds2 / 2.0 - ds3 / 3.0)) / c;
}

// the vertical part of calc_grad_ptf (operators_mod.F90:1083-1101): the
// chunk's cells k0 .. k1-1 need their faces k0 .. k1; the interior ones are
// formed here (face k0 also by the chunk above, which writes it), the top
// and bottom faces are read as they are.  flag[0] is set where a Courant
// number reaches outside the column.  LIM: the limiter, whose branches fold
// at compile time.
template <int LIM>
__global__ void __launch_bounds__(NB) k_vflx_pt(BaroGrid g, const double *__restrict__ pt,
                                                const double *__restrict__ we, const double *__restrict__ cflz,
                                                double *ptf, double *dptf, int *flag) {
  constexpr int lim = LIM;
  int64_t o;
  int i, r, ch;
  if (!plane_point(g, g.nlat, o, i, r, ch)) return;
  const int64_t s = g.sc;
  const int n = g.nlev, k0 = ch * LC, k1 = min(k0 + LC, n);
  const double *col = pt + o;
  bool bad = false;
  double top = k0 == 0 ? ptf[o] : pt_face_flux(lim, col, s, n, k0, cflz[k0 * s + o], we[k0 * s + o], bad);
  for (int k = k0; k < k1; k++) {
    double bot;
    if (k + 1 == n) {
      bot = ptf[n * s + o];
    } else {
      bot = pt_face_flux(lim, col, s, n, k + 1, cflz[(k + 1) * s + o], we[(k + 1) * s + o], bad);
      ptf[(k + 1) * s + o] = bot;
    }
    dptf[k * s + o] = bot - top;
    top = bot;
  }
  if (bad) atomicOr(flag, kFlagCflz);
}

// The horizontal part of calc_grad_ptf (operators_mod.F90:1029-1080):
// ffsl_calc_tracer_hflx_cell (ffsl_mod.F90:318-413) of pt with the mass
// batch's fields on every level, then the flux divergence.  A thread owns one
// point of one level (grid y: the level); the face fluxes are those of the
// tracer advection (gmc_hflx.hpp).

// x-fluxes of mx through the U points of rows 1 .. nlat-2 and y-fluxes of my
// through every V point, with the winds u / v and Courant numbers cflx / cfly
// (hflx_ppm_cell / hflx_van_leer_cell, ffsl_mod.F90:556-616, :778-881): the
// thread of cell (i, r), r = 0 .. nlat-2, forms the V point north of it and,
// off the south pole row, the U point east of it.  A cflx beyond
// lon_halo_width sets kFlagCflx (adv_batch_mod.F90:284-287); its whole-cell
// count is clamped to the bound so that every access stays in the row.
template <bool VL, int LIM>
__global__ void __launch_bounds__(NB) k_hflx_levels(BaroGrid g, AdvGrid ag, const double *__restrict__ u,
                                                    const double *__restrict__ v, const double *__restrict__ cflx,
                                                    const double *__restrict__ cfly, const double *__restrict__ mx,
                                                    const double *__restrict__ my, double *fx, double *fy, int *flag) {
  constexpr int lim = LIM;
  int64_t o;
  int i, r, k;
  if (!plane_point<kByRows>(g, g.nlat - 1, o, i, r, k)) return;
  const int64_t c = k * g.sc + o, cv = k * g.sv + o;
  if (r >= 1) {
    const double cx = cflx[c];
    if (!(fabs(cx) <= ag.hw)) atomicOr(flag, kFlagCflx);
    fx[c] = xflux<VL>(ag, lim, cx, u[c], i, mx + (c - i));
  }
  fy[cv] = yflux_any<VL>(ag, lim, cfly[cv], v[cv], r, i, my + k * g.sc);
}

using VflxKernel = void (*)(BaroGrid, const double *, const double *, const double *, double *, double *, int *);
VflxKernel vflx_kernel(int lim) {
  return lim == GMC_LIMITER_PD ? k_vflx_pt<GMC_LIMITER_PD>
       : lim == GMC_LIMITER_NONE ? k_vflx_pt<GMC_LIMITER_NONE> : k_vflx_pt<GMC_LIMITER_MONO>;
}
using HflxKernel = void (*)(BaroGrid, AdvGrid, const double *, const double *, const double *, const double *,
                            const double *, const double *, double *, double *, int *);
template <bool VL>
HflxKernel hflx_kernel_lim(int lim) {
  return lim == GMC_LIMITER_PD ? k_hflx_levels<VL, GMC_LIMITER_PD>
       : lim == GMC_LIMITER_NONE ? k_hflx_levels<VL, GMC_LIMITER_NONE> : k_hflx_levels<VL, GMC_LIMITER_MONO>;
}
HflxKernel hflx_kernel(bool vl, int lim) { return vl ? hflx_kernel_lim<true>(lim) : hflx_kernel_lim<false>(lim); }

// the advective half-steps qx, qy of the cells off the poles
// (ffsl_mod.F90:354-373) from the inner fluxes qmfx, qmfy
__global__ void __launch_bounds__(NB) k_half_steps(BaroGrid g, Mesh M, double dt, const double *__restrict__ q,
                                                   const double *__restrict__ qmfx, const double *__restrict__ qmfy,
                                                   const double *__restrict__ divx, const double *__restrict__ divy,
                                                   double *qx, double *qy) {
  int64_t o;
  int i, r, k;
  if (!plane_point<kByRows>(g, g.nlat, o, i, r, k) || r == 0 || r == g.nlat - 1) return;
  const int64_t c = k * g.sc + o, cv = k * g.sv + o;
  const int iw = i == 0 ? g.nlon - 1 : i - 1;
  const double qc = q[c];
  const double fx = qmfx[c], fxw = qmfx[c - i + iw], fn = qmfy[cv], fs = qmfy[cv - g.pitch];
  const double lel = M(GMC_LE_LON, r), ac = M(GMC_AREA_CELL, r), ln = M(GMC_LE_LAT, r), ls = M(GMC_LE_LAT, r - 1);
  qx[c] = qc - 0.5 * ((fx - fxw) * lel / ac - divx[c] * qc) * dt;
  qy[c] = qc - 0.5 * ((fn * ln - fs * ls) / ac - divy[c] * qc) * dt;
}

// the pole rows' half-steps (ffsl_mod.F90:374-406): qx = q, qy from the
// zonal sum of the adjacent V row's inner flux; one wavefront per (pole,
// level), the sum in the library's fixed order
__global__ void __launch_bounds__(kWave) k_half_steps_pole(BaroGrid g, Mesh M, double dt,
                                                           const double *__restrict__ q,
                                                           const double *__restrict__ qmfy,
                                                           const double *__restrict__ divy, double *qx, double *qy) {
  const bool north = blockIdx.x == 1;
  const int r = north ? g.nlat - 1 : 0, h = north ? g.nlat - 2 : 0;
  const int64_t k = blockIdx.y;
  const double *fy = qmfy + k * g.sv + h * g.pitch;
  double pole = wave_zonal_sum(g.nlon, threadIdx.x, [&](int i) { return fy[i]; });
  pole = pole * M(GMC_LE_LAT, h) / g.nlon / M(GMC_AREA_CELL, r);
  for (int i = threadIdx.x; i < g.nlon; i += kWave) {
    const int64_t c = k * g.sc + r * g.pitch + i;
    const double qc = q[c];
    qx[c] = qc;
    qy[c] = north ? qc + 0.5 * (pole - divy[c] * qc) * dt : qc - 0.5 * (pole - divy[c] * qc) * dt;
  }
}

// dptfdlon and dptfdlat of the cells off the poles (operators_mod.F90:1034-1051)
__global__ void __launch_bounds__(NB) k_ptf_div(BaroGrid g, Mesh M, const double *__restrict__ fxl,
                                                const double *__restrict__ fyl, double *dx, double *dy) {
  int64_t o;
  int i, r, k;
  if (!plane_point<kByRows>(g, g.nlat, o, i, r, k) || r == 0 || r == g.nlat - 1) return;
  const int64_t c = k * g.sc + o, cv = k * g.sv + o;
  const int iw = i == 0 ? g.nlon - 1 : i - 1;
  const double ac = M(GMC_AREA_CELL, r);
  dx[c] = (fxl[c] - fxl[c - i + iw]) * M(GMC_LE_LON, r) / ac;
  dy[c] = (fyl[cv] * M(GMC_LE_LAT, r) - fyl[cv - g.pitch] * M(GMC_LE_LAT, r - 1)) / ac;
}

// dptfdlat of the pole rows (operators_mod.F90:1052-1079): the zonal sum of
// the adjacent ptf_lat row (negated at the north pole), one wavefront per
// (pole, level)
__global__ void __launch_bounds__(kWave) k_ptf_div_pole(BaroGrid g, Mesh M, const double *__restrict__ fyl,
                                                        double *dy) {
  const bool north = blockIdx.x == 1;
  const int r = north ? g.nlat - 1 : 0, h = north ? g.nlat - 2 : 0;
  const int64_t k = blockIdx.y;
  const double *fy = fyl + k * g.sv + h * g.pitch;
  double pole = wave_zonal_sum(g.nlon, threadIdx.x, [&](int i) { return north ? -fy[i] : fy[i]; });
  pole = pole * M(GMC_LE_LAT, h) / g.nlon / M(GMC_AREA_CELL, r);
  for (int i = threadIdx.x; i < g.nlon; i += kWave) dy[k * g.sc + r * g.pitch + i] = pole;
}

// filter_on_lon_edge / filter_on_lat_edge (filter_mod.F90:93-163) in place:
// block (row, level) stages its row and the halo of its taps in LDS, then
// writes the filtered row (filter_point: products and additions in tap order)
__global__ void __launch_bounds__(NB) k_filter_levels(BaroGrid g, int rows, int64_t lstride,
                                                      const int32_t *__restrict__ ngrid,
                                                      const double *__restrict__ wgt, int max_taps, double *x) {
  extern __shared__ double row[];
  const int r = blockIdx.x, k = blockIdx.y;
  if (r >= rows) return;
  const int n = ngrid[r];
  if (n <= 1) return;
  const int hn = (n - 1) / 2, len = g.nlon + 2 * hn;
  double *xr = x + k * lstride + r * g.pitch;
  for (int t = threadIdx.x; t < len; t += NB) {
    int c = (t - hn) % g.nlon;
    if (c < 0) c += g.nlon;
    row[t] = xr[c];
  }
  __syncthreads();
  const double *w = wgt + (int64_t)r * max_taps;
  for (int i = threadIdx.x; i < g.nlon; i += NB) xr[i] = filter_point(w, n, row + i);
}

}  // namespace
}  // namespace gmc

using namespace gmc;

struct gmc_baro {
  BaroGrid g{};
  double vp0 = 0, ptop = 0;          // hybrid_coord_p0, local_ptop
  double Rd = 0, rd_o_cpd = 0, tp0 = 0;
  double *slab = nullptr;
  double *f[GMB_NFIELDS] = {};
  double *d_mesh = nullptr, *d_hyai = nullptr, *d_hybi = nullptr;
  double *w_qmfx = nullptr, *w_qmfy = nullptr, *w_qx = nullptr, *w_qy = nullptr;   // calc_ptf_hflx work
  int *d_flag = nullptr;             // error bits of calc_ptf_lev / calc_ptf_hflx
  bool flag_pending = false;         // one of them launched since the flag was last read
  void *d_filter = nullptr;          // small_filter2's table on the device
  size_t filter_bytes = 0;
  int32_t *f_ngrid = nullptr;        // its tap counts: U rows, then V rows
  const double *f_wlon = nullptr, *f_wlat = nullptr;
  int f_max_taps = 0, f_hmax = 0;    // 0: no table set
  hipStream_t stream = nullptr;
  HostStage stage;
  std::string err;
};

namespace {

int berr(gmc_baro *b, int code, const std::string &msg) {
  if (b) b->err = msg;
  return code;
}

int bhip(gmc_baro *b, hipError_t e, const char *what) {
  if (e == hipSuccess) return GMC_OK;
  return berr(b, e == hipErrorOutOfMemory ? GMC_ENOMEM : GMC_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int levels_of(const gmc_baro *b, int f) {
  const int l = kShape[f].lev;
  return l == 0 ? 1 : (l == 1 ? b->g.nlev : b->g.nlev + 1);
}
int rows_of(const gmc_baro *b, int f) { return kShape[f].vrows ? b->g.nlat - 1 : b->g.nlat; }

dim3 columns(const BaroGrid &g, int rows) { return dim3((g.nlon + NB - 1) / NB, rows); }

// a level-chunked launch: every point of `rows` rows of a level plane, the
// levels in chunks of LC
dim3 chunks(const BaroGrid &g, int rows, int levels, int lc = LC) {
  return dim3((unsigned)(((int64_t)rows * g.pitch + NB - 1) / NB), (levels + lc - 1) / lc);
}

// calc_gz_lev's columns per block: the widest power of two up to 64 whose
// layer terms and half levels fit 64 KiB of LDS
int gz_columns(int nlev) {
  int cw = 64;
  while (cw > 1 && 2 * (size_t)nlev * cw * sizeof(double) > 65536) cw /= 2;
  return cw;
}

int launched(gmc_baro *b, const char *what) { return bhip(b, hipGetLastError(), what); }

}  // namespace

extern "C" {

int gmc_baro_create(const gmc_mesh *mesh, const gmc_vcoord *vc, const gmc_thermo *th, gmc_baro **out) {
  if (!out) return GMC_EINVAL;
  *out = nullptr;
  if (!mesh || !vc || !th || !mesh->metrics || !vc->hyai || !vc->hybi) return GMC_EINVAL;
  const int nlon = mesh->nlon, nlat = mesh->nlat, nlev = vc->nlev;
  if (nlon < 8 || nlat < 5 || (nlon & 1) || nlev < 2 || nlev > 4096 || !(th->cpd > 0) || !(th->p0 > 0) ||
      !(vc->p0 > 0))
    return GMC_EINVAL;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return GMC_EDEVICE;
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0) return GMC_EDEVICE;
  gmc_baro *b = new gmc_baro();
  BaroGrid &g = b->g;
  g.nlon = nlon;
  g.nlat = nlat;
  g.nlev = nlev;
  g.pitch = (nlon + 31) / 32 * 32;
  g.sc = (int64_t)nlat * g.pitch;
  g.sv = (int64_t)(nlat - 1) * g.pitch;
  b->vp0 = vc->p0;
  b->ptop = vc->ptop;
  b->Rd = th->Rd;
  b->rd_o_cpd = th->Rd / th->cpd;   // Rd_o_cpd (const_mod.F90:81)
  b->tp0 = th->p0;
  size_t n = 0;
  for (int f = 0; f < GMB_NFIELDS; f++) n += (size_t)levels_of(b, f) * rows_of(b, f) * g.pitch;
  n += GMC_MESH_NFIELDS * (size_t)nlat + 2 * (size_t)(nlev + 1);
  n += (size_t)nlev * (3 * g.sc + g.sv);          // calc_ptf_hflx's work: qmfx, qx, qy; qmfy
  n += 4;                                          // the error flag
  int rc = bhip(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
  if (!rc) rc = bhip(b, hipMalloc((void **)&b->slab, n * sizeof(double)), "alloc");
  if (!rc) rc = bhip(b, hipMemset(b->slab, 0, n * sizeof(double)), "memset");
  if (rc) {
    gmc_baro_destroy(b);
    return rc;
  }
  double *p = b->slab;
  for (int f = 0; f < GMB_NFIELDS; f++) {
    b->f[f] = p;
    p += (size_t)levels_of(b, f) * rows_of(b, f) * g.pitch;
  }
  b->d_mesh = p;
  p += GMC_MESH_NFIELDS * (size_t)nlat;
  b->d_hyai = p;
  b->d_hybi = p + nlev + 1;
  p += 2 * (size_t)(nlev + 1);
  b->w_qmfx = p;
  p += (size_t)nlev * g.sc;
  b->w_qx = p;
  p += (size_t)nlev * g.sc;
  b->w_qy = p;
  p += (size_t)nlev * g.sc;
  b->w_qmfy = p;
  p += (size_t)nlev * g.sv;
  b->d_flag = (int *)p;
  rc = bhip(b, hipMemcpy(b->d_mesh, mesh->metrics, sizeof(double) * GMC_MESH_NFIELDS * nlat, hipMemcpyHostToDevice),
            "mesh");
  if (!rc) rc = bhip(b, hipMemcpy(b->d_hyai, vc->hyai, sizeof(double) * (nlev + 1), hipMemcpyHostToDevice), "hyai");
  if (!rc) rc = bhip(b, hipMemcpy(b->d_hybi, vc->hybi, sizeof(double) * (nlev + 1), hipMemcpyHostToDevice), "hybi");
  if (rc) {
    gmc_baro_destroy(b);
    return rc;
  }
  *out = b;
  return GMC_OK;
}

int gmc_baro_destroy(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  if (b->stream) hip_note(hipStreamSynchronize(b->stream), "baro sync");
  if (b->slab) hip_note(hipFree(b->slab), "baro free");
  if (b->d_filter) hip_note(hipFree(b->d_filter), "baro filter free");
  stage_free(b->stage);
  if (b->stream) hip_note(hipStreamDestroy(b->stream), "baro stream");
  delete b;
  return GMC_OK;
}

const char *gmc_baro_last_error(const gmc_baro *b) { return b ? b->err.c_str() : "null object"; }

int gmc_baro_set(gmc_baro *b, int32_t field, const double *host, int64_t ld) {
  if (!b) return GMC_EINVAL;
  if (field < 0 || field >= GMB_NFIELDS || !host || ld < b->g.nlon) return berr(b, GMC_EINVAL, "bad field upload");
  return bhip(b, stage_copy(b->stage, b->stream, true, b->f[field], b->g.pitch, const_cast<double *>(host), ld,
                            b->g.nlon, levels_of(b, field) * rows_of(b, field)),
              "upload");
}

int gmc_baro_get(gmc_baro *b, int32_t field, double *host, int64_t ld) {
  if (!b) return GMC_EINVAL;
  if (field < 0 || field >= GMB_NFIELDS || !host || ld < b->g.nlon) return berr(b, GMC_EINVAL, "bad field download");
  return bhip(b, stage_copy(b->stage, b->stream, false, b->f[field], b->g.pitch, host, ld, b->g.nlon,
                            levels_of(b, field) * rows_of(b, field)),
              "download");
}

int gmc_baro_calc_ph(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(k_calc_ph, chunks(g, g.nlat, g.nlev, LCP), dim3(NB), 0, b->stream, g, b->d_hyai, b->d_hybi, b->vp0, b->ptop,
                     b->rd_o_cpd, b->f[GMB_PHS], b->f[GMB_PH_LEV], b->f[GMB_PH_EXN_LEV], b->f[GMB_PH]);
  return launched(b, "calc_ph");
}

int gmc_baro_calc_t(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(k_calc_t, chunks(g, g.nlat, g.nlev, LCP), dim3(NB), 0, b->stream, g, b->tp0, b->rd_o_cpd, b->f[GMB_PT],
                     b->f[GMB_PH], b->f[GMB_T]);
  return launched(b, "calc_t");
}

int gmc_baro_calc_gz_lev(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  const int cw = gz_columns(g.nlev);
  const unsigned blocks = (unsigned)(((int64_t)g.nlat * g.pitch + cw - 1) / cw);
  hipLaunchKernelGGL(k_calc_gz_lev, dim3(blocks), dim3(NB), 2 * (size_t)g.nlev * cw * sizeof(double), b->stream, g,
                     cw, b->Rd, b->f[GMB_GZS], b->f[GMB_T], b->f[GMB_PH_LEV], b->f[GMB_GZ_LEV], b->f[GMB_GZ]);
  return launched(b, "calc_gz_lev");
}

int gmc_baro_calc_dphsdt(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(k_calc_dphsdt, columns(g, g.nlat), dim3(NB), 0, b->stream, g, b->f[GMB_DMFDLON],
                     b->f[GMB_DMFDLAT], b->f[GMB_DPHS]);
  return launched(b, "calc_dphsdt");
}

int gmc_baro_calc_we_lev(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(k_calc_we_lev, columns(g, g.nlat), dim3(NB), 0, b->stream, g, b->d_hybi, b->f[GMB_DMFDLON],
                     b->f[GMB_DMFDLAT], b->f[GMB_DPHS], b->f[GMB_WE_LEV]);
  hipLaunchKernelGGL(k_we_edges, chunks(g, g.nlat - 1, g.nlev + 1), dim3(NB), 0, b->stream, g, Mesh{b->d_mesh, g.nlat},
                     b->f[GMB_WE_LEV], b->f[GMB_WE_LEV_LON], b->f[GMB_WE_LEV_LAT]);
  return launched(b, "calc_we_lev");
}

int gmc_baro_calc_wedudlev_wedvdlev(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(k_wedudlev, chunks(g, 2 * g.nlat - 3, g.nlev), dim3(NB), 0, b->stream, g, b->f[GMB_U_F],
                     b->f[GMB_WE_LEV_LON], b->f[GMB_M_LON], b->f[GMB_WEDUDLEV], b->f[GMB_V_F], b->f[GMB_WE_LEV_LAT],
                     b->f[GMB_M_LAT], b->f[GMB_WEDVDLEV]);
  return launched(b, "calc_wedudlev_wedvdlev");
}

int gmc_baro_pgf_lin97(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(k_pgf_lin97, chunks(g, g.nlat - 1, g.nlev), dim3(NB), 0, b->stream, g, Mesh{b->d_mesh, g.nlat},
                     b->f[GMB_PH_EXN_LEV], b->f[GMB_GZ_LEV], b->f[GMB_PGF_LON], b->f[GMB_PGF_LAT]);
  return launched(b, "pgf_lin97");
}

int gmc_baro_calc_ptf_lev(gmc_baro *b, int32_t limiter) {
  if (!b) return GMC_EINVAL;
  if (limiter < GMC_LIMITER_MONO || limiter > GMC_LIMITER_NONE) return berr(b, GMC_EINVAL, "bad limiter");
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(vflx_kernel(limiter), chunks(g, g.nlat, g.nlev), dim3(NB), 0, b->stream, g, b->f[GMB_PT],
                     b->f[GMB_WE_LEV], b->f[GMB_CFLZ], b->f[GMB_PTF_LEV], b->f[GMB_DPTFDLEV], b->d_flag);
  const int rc = launched(b, "calc_ptf_lev");
  if (!rc) b->flag_pending = true;
  return rc;
}

int gmc_baro_calc_ptf_hflx(gmc_baro *b, int32_t limiter, int32_t flux_type, int32_t lon_halo_width, double dt) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  if (limiter < GMC_LIMITER_MONO || limiter > GMC_LIMITER_NONE) return berr(b, GMC_EINVAL, "bad limiter");
  if (flux_type != GMC_FLUX_PPM && flux_type != GMC_FLUX_VAN_LEER) return berr(b, GMC_EINVAL, "bad flux type");
  if (lon_halo_width < 1 || 2 * lon_halo_width >= g.nlon) return berr(b, GMC_EINVAL, "bad lon_halo_width");
  if (!std::isfinite(dt)) return berr(b, GMC_EINVAL, "bad dt");
  AdvGrid ag{};
  ag.nlon = g.nlon;
  ag.nlat = g.nlat;
  ag.nhalf = g.nlat - 1;
  ag.half = g.nlon / 2;
  ag.pitch = g.pitch;
  ag.hw = lon_halo_width;
  ag.dt = dt;
  const Mesh M{b->d_mesh, g.nlat};
  const dim3 faces = dim3(chunks(g, g.nlat - 1, 1).x, g.nlev), cells = dim3(chunks(g, g.nlat, 1).x, g.nlev);
  const HflxKernel hk = hflx_kernel(flux_type == GMC_FLUX_VAN_LEER, limiter);
  double **f = b->f;
  // inner fluxes of pt with the winds, then the half-steps
  hipLaunchKernelGGL(hk, faces, dim3(NB), 0, b->stream, g, ag, f[GMB_ADV_U], f[GMB_ADV_V], f[GMB_ADV_CFLX],
                     f[GMB_ADV_CFLY], f[GMB_PT], f[GMB_PT], b->w_qmfx, b->w_qmfy, b->d_flag);
  hipLaunchKernelGGL(k_half_steps, cells, dim3(NB), 0, b->stream, g, M, dt, f[GMB_PT], b->w_qmfx, b->w_qmfy,
                     f[GMB_ADV_DIVX], f[GMB_ADV_DIVY], b->w_qx, b->w_qy);
  hipLaunchKernelGGL(k_half_steps_pole, dim3(2, g.nlev), dim3(kWave), 0, b->stream, g, M, dt, f[GMB_PT], b->w_qmfy,
                     f[GMB_ADV_DIVY], b->w_qx, b->w_qy);
  // outer fluxes of qy along x and qx along y with the mass fluxes
  hipLaunchKernelGGL(hk, faces, dim3(NB), 0, b->stream, g, ag, f[GMB_ADV_MFX], f[GMB_ADV_MFY], f[GMB_ADV_CFLX],
                     f[GMB_ADV_CFLY], b->w_qy, b->w_qx, f[GMB_PTF_LON], f[GMB_PTF_LAT], b->d_flag);
  hipLaunchKernelGGL(k_ptf_div, cells, dim3(NB), 0, b->stream, g, M, f[GMB_PTF_LON], f[GMB_PTF_LAT],
                     f[GMB_DPTFDLON], f[GMB_DPTFDLAT]);
  hipLaunchKernelGGL(k_ptf_div_pole, dim3(2, g.nlev), dim3(kWave), 0, b->stream, g, M, f[GMB_PTF_LAT],
                     f[GMB_DPTFDLAT]);
  const int rc = launched(b, "calc_ptf_hflx");
  if (!rc) b->flag_pending = true;
  return rc;
}

int gmc_baro_calc_grad_ptf(gmc_baro *b, int32_t limiter, int32_t flux_type, int32_t lon_halo_width, double dt) {
  int rc = gmc_baro_calc_ptf_hflx(b, limiter, flux_type, lon_halo_width, dt);
  if (!rc) rc = gmc_baro_calc_ptf_lev(b, limiter);
  return rc;
}

int gmc_baro_set_small_filter2(gmc_baro *b, const gmc_filter *f) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  const int nlat = g.nlat, nhalf = nlat - 1;
  if (!f || !f->ngrid_lon || !f->ngrid_lat || !f->wgt_lon || !f->wgt_lat || f->max_taps < 1)
    return berr(b, GMC_EINVAL, "bad filter table");
  int hmax = 0;
  for (int r = 0; r < nlat; r++) {
    if (f->ngrid_lon[r] > g.nlon || f->ngrid_lon[r] > f->max_taps) return berr(b, GMC_EINVAL, "bad filter taps");
    hmax = std::max(hmax, (f->ngrid_lon[r] - 1) / 2);
  }
  for (int r = 0; r < nhalf; r++) {
    if (f->ngrid_lat[r] > g.nlon || f->ngrid_lat[r] > f->max_taps) return berr(b, GMC_EINVAL, "bad filter taps");
    hmax = std::max(hmax, (f->ngrid_lat[r] - 1) / 2);
  }
  if (sizeof(double) * (g.nlon + 2 * (size_t)hmax) > 65536) return berr(b, GMC_EINVAL, "filter row does not fit in LDS");
  const size_t ni = sizeof(int32_t) * (nlat + nhalf), nw = sizeof(double) * (size_t)f->max_taps * (nlat + nhalf);
  const size_t need = (ni + 255) / 256 * 256 + nw;
  // the launches that read the old table finish first
  int rc = bhip(b, hipStreamSynchronize(b->stream), "sync");
  if (rc) return rc;
  b->f_max_taps = 0;
  if (need > b->filter_bytes) {
    if (b->d_filter) rc = bhip(b, hipFree(b->d_filter), "filter free");
    b->d_filter = nullptr;
    b->filter_bytes = 0;
    if (!rc) rc = bhip(b, hipMalloc(&b->d_filter, need), "filter alloc");
    if (rc) return rc;
    b->filter_bytes = need;
  }
  int32_t *ng = (int32_t *)b->d_filter;
  double *w = (double *)((char *)b->d_filter + (ni + 255) / 256 * 256);
  const size_t wl = sizeof(double) * (size_t)f->max_taps * nlat;
  rc = bhip(b, hipMemcpy(ng, f->ngrid_lon, sizeof(int32_t) * nlat, hipMemcpyHostToDevice), "filter copy");
  if (!rc) rc = bhip(b, hipMemcpy(ng + nlat, f->ngrid_lat, sizeof(int32_t) * nhalf, hipMemcpyHostToDevice), "filter copy");
  if (!rc) rc = bhip(b, hipMemcpy(w, f->wgt_lon, wl, hipMemcpyHostToDevice), "filter copy");
  if (!rc) rc = bhip(b, hipMemcpy((char *)w + wl, f->wgt_lat, sizeof(double) * (size_t)f->max_taps * nhalf,
                                  hipMemcpyHostToDevice), "filter copy");
  if (rc) return rc;
  b->f_ngrid = ng;
  b->f_wlon = w;
  b->f_wlat = (const double *)((const char *)w + wl);
  b->f_hmax = hmax;
  b->f_max_taps = f->max_taps;
  return GMC_OK;
}

int gmc_baro_small_filter2(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  if (!b->f_max_taps) return berr(b, GMC_EINVAL, "no small_filter2 table set");
  const BaroGrid &g = b->g;
  const size_t lds = sizeof(double) * (g.nlon + 2 * (size_t)b->f_hmax);
  hipLaunchKernelGGL(k_filter_levels, dim3(g.nlat, g.nlev), dim3(NB), lds, b->stream, g, g.nlat, g.sc, b->f_ngrid,
                     b->f_wlon, b->f_max_taps, b->f[GMB_U_LON]);
  hipLaunchKernelGGL(k_filter_levels, dim3(g.nlat - 1, g.nlev), dim3(NB), lds, b->stream, g, g.nlat - 1, g.sv,
                     b->f_ngrid + g.nlat, b->f_wlat, b->f_max_taps, b->f[GMB_V_LAT]);
  return launched(b, "small_filter2");
}

int gmc_baro_synchronize(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  int rc = bhip(b, hipStreamSynchronize(b->stream), "synchronize");
  if (rc || !b->flag_pending) return rc;
  int h = 0;
  rc = bhip(b, hipMemcpy(&h, b->d_flag, sizeof(int), hipMemcpyDeviceToHost), "flag");
  if (rc) return rc;
  b->flag_pending = false;
  if (!h) return GMC_OK;
  const int zero = 0;
  rc = bhip(b, hipMemcpy(b->d_flag, &zero, sizeof(int), hipMemcpyHostToDevice), "flag clear");
  if (rc) return rc;
  std::string msg;
  if (h & kFlagCflz) msg = "cflz reaches outside the column";
  if (h & kFlagCflx) msg += std::string(msg.empty() ? "" : "; ") + "cflx exceeds lon_halo_width";
  return berr(b, GMC_ENUMERIC, msg);
}

int gmc_baro_stream(gmc_baro *b, void **stream) {
  if (!b || !stream) return GMC_EINVAL;
  *stream = (void *)b->stream;
  return GMC_OK;
}

}  // extern "C"
