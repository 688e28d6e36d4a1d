// gmc_fused.hip -- fused stage kernels used by gmc_step / gmc_time_integrate.
//
// One step_forward_backward stage (time_schemes_mod.F90:102-118) of the SW
// path is three launches:
//   k_tend   an LDS-tiled stencil over 58 x 11 cells: the pole sums (in the
//            tiles next to the poles), m, mass fluxes, pv,
//            ke, tangential winds, upwind pv on edges, then the forward
//            tendencies du, dv and the updated, unfiltered gz
//            (operators_prepare_2 forward + calc_grad_mf + calc_coriolis +
//            calc_grad_ke + the gz part of update_state);
//   k_gzf    the big and small zonal filters of the new gz;
//   k_uvf    the backward pass (pgf from gz_f), the u/v update with the
//            pole-row smoothing, and the big filters to u_f, v_f.
// The filter kernels run over a host-built work list of (row, column chunk)
// items so that polar rows with hundreds of taps are spread over many
// workgroups.  Every expression is the same point formula as the
// operator-level kernels (gmc_point.hpp), so both paths give the same bits.
#include <hip/hip_ext.h>

#include <algorithm>
#include <vector>

#include "gmc_internal.hpp"
#include "gmc_point.hpp"
#include "../../include/gmcore_ens.h"

namespace gmc {

namespace {

// k_tend tile: one wavefront per tile row, lane q <-> column C0 - 3 + q
constexpr int LC = kWave;               // 64 tile columns
constexpr int HXL = 3;                  // column halo
constexpr int TX = LC - 2 * HXL - 0;    // 58 output columns (q = 3 .. 60)
// launch-shape parameters (the defaults are the measured best on MI355X;
// the macros let a sweep build other shapes)
#ifndef GMC_RPW_MAIN
#define GMC_RPW_MAIN 1
#endif
#ifndef GMC_TY_MAIN
#define GMC_TY_MAIN (GMC_RPW_MAIN == 2 ? 27 : 11)
#endif
#ifndef GMC_FILTER_THREADS
#define GMC_FILTER_THREADS 256
#endif
#ifndef GMC_CHUNK_WORK
#define GMC_CHUNK_WORK 32768
#endif
#ifndef GMC_NARROW_CHUNK
#define GMC_NARROW_CHUNK (2 * GMC_FILTER_THREADS)
#endif
constexpr int TY_MAIN = GMC_TY_MAIN;    // output rows of the single-member k_tend tile
constexpr int RPW_MAIN = GMC_RPW_MAIN;  // its tile rows per wavefront
constexpr int NF = GMC_FILTER_THREADS;  // threads per workgroup of the filter kernels
constexpr int NXCD = 8;                 // XCDs of an MI355X
// A k_tend tile with TY output rows: LR = TY + 5 tile rows (r = R0-2 ..
// R0+TY+2), RPW tile rows per wavefront: wavefront w owns the tile rows
// w, w + NW, .. (NW = LR / RPW wavefronts).
template <int TY_, int RPW_ = 1> struct Tile {
  static constexpr int TY = TY_;
  static constexpr int RPW = RPW_;
  static constexpr int LR = TY + 5;
  static_assert(LR % RPW == 0, "the tile rows split evenly over the wavefronts");
  static constexpr int NW = LR / RPW;
  static constexpr int NT = NW * kWave;
  static constexpr int NE = LR * LC;
  static constexpr int SH_TILES = 7 * NE;                 // LDS field tiles
  static constexpr int SH_MESH = LR * kRowRec;            // metric row records of the tile
  static constexpr int NREC = (SH_MESH + NT - 1) / NT;    // record words per thread
  static constexpr int SH_ALL = SH_TILES + SH_MESH + 8;
  static constexpr int kMaxPoleLon = SH_TILES / 3;        // the pole sums keep 3 row terms in the tile area
};

struct StagePtrs {
  const double *gz_star, *gzs, *u_f, *v_f, *gz_old;
  double *gz_new, *du, *dv;
};

// the `order` argument of pv_edge: upwind_order_pv, or the limiter for tvd
inline int pv_order(const gmc_config &cfg) {
  return cfg.pv_scheme == GMC_PV_TVD ? cfg.tvd_limiter : cfg.upwind_order_pv;
}

// pv at vertex (r, i) from global memory: rows reached through the
// cross-pole mirror next to the poles
template <class MT, class PT>
__device__ double pv_global(const Grid &g, const MT &M, double grav, double rgrav, int stokes, const double *pole,
                            const PT &P, int r, int i) {
  const int64_t p = g.pitch;
  const int ip = wrap(i + 1, g.nlon);
  auto m = [&](int rr, int ii) { return div_rn(P.gz_star[rr * p + ii] - P.gzs[rr * p + ii], grav, rgrav); };
  double vor;
  if (stokes && r == 0) vor = pole[4];
  else if (stokes && r == g.nhalf - 1) vor = pole[5];
  else vor = vor_of(P.u_f[r * p + i], P.u_f[(r + 1) * p + i], P.v_f[r * p + ip], P.v_f[r * p + i], M, r);
  return pv_of(vor, M, r, m_vtx_of(m(r, i), m(r, ip), m(r + 1, i), m(r + 1, ip), M, r));
}

// Pole reductions of the star state for one pole (zonal_sum call sites in
// calc_ke :320-349, calc_vor :640-674, calc_grad_mf :974-1003): the row
// terms go to LDS scratch, one wavefront adds them in the fixed order of
// wave_zonal_sum.  pole[north] = ke, pole[2+north] = dmfdlat at the pole
// cell, pole[4+north] = vor of the vertex row next to the pole.
template <int NT, class MT, class PT>
__device__ void pole_sums(const Grid &g, const MT &M, double grav, double rgrav, const PT &P, int north,
                          double *scratch, double *pole) {
  const int tid = threadIdx.x, n = g.nlon;
  const int64_t p = g.pitch;
  const int vr = north ? g.nhalf - 1 : 0, cr = north ? g.nlat - 1 : 0, ur = north ? g.nlat - 2 : 1;
  const double le = M(GMC_LE_LAT, vr), ac = M(GMC_AREA_CELL, cr);
  const double w1 = M(GMC_HALF_TANGENT_WGT1, vr), w2 = M(GMC_HALF_TANGENT_WGT2, vr);
  double *t0 = scratch, *t1 = scratch + n, *t2 = scratch + 2 * n;
  auto m = [&](int r, int i) { return div_rn(P.gz_star[r * p + i] - P.gzs[r * p + i], grav, rgrav); };
  for (int i = tid; i < n; i += NT) {
    const int im = wrap(i - 1, n), ip = wrap(i + 1, n);
    const double v = P.v_f[vr * p + i];
    const double m_lat = (m(vr, i) + m(vr + 1, i)) * 0.5;
    const double mu = m(ur, i);
    const double mfx_w = (m(ur, im) + mu) * 0.5 * P.u_f[ur * p + im];
    const double mfx_e = (mu + m(ur, ip)) * 0.5 * P.u_f[ur * p + i];
    // mfx_lat of the V row: the pole U row contributes zero flux
    const double t = north ? tangent_sum(w1, w2, mfx_w, mfx_e, 0.0, 0.0) : tangent_sum(w1, w2, 0.0, 0.0, mfx_w, mfx_e);
    const double ul = t / m_lat;
    const double mfy = m_lat * v;
    t0[i] = v * v;
    t1[i] = north ? -mfy : mfy;
    t2[i] = north ? ul * le : -ul * le;
  }
  __syncthreads();
  if (tid < kWave) {
    const double ske = wave_zonal_sum(n, tid, [&](int i) { return t0[i]; });
    const double smf = wave_zonal_sum(n, tid, [&](int i) { return t1[i]; });
    const double svo = wave_zonal_sum(n, tid, [&](int i) { return t2[i]; });
    if (tid == 0) {
      const double nl = (double)n;
      pole[north] = ske / nl;
      pole[2 + north] = smf * le / nl / ac;
      pole[4 + north] = svo / nl / ac;
    }
  }
  __syncthreads();
}

// Metric rows of one tile in LDS (row records, tile row k = r - r0):
// M(f, r) is a broadcast LDS read for a wave-uniform r.  The records carry
// the reciprocals of the divisor fields, so M.div divides through div_rn.
struct MeshTile {
  const double *s;
  int r0;
  __device__ __forceinline__ double operator()(int f, int r) const { return s[(r - r0) * kRowRec + f]; }
  __device__ __forceinline__ double div(double x, int f, int r) const {
    const double *rec = s + (r - r0) * kRowRec;
    return div_rn(x, rec[f], rec[GMC_MESH_NFIELDS + recip_slot(f)]);
  }
};

// The global loads of one tile, per lane and tile row of its wavefront: gz,
// gzs, u_f, v_f of its tile point and gz_old of its output point; and its
// words of the metric records.
template <class S> struct TileIn {
  double gz[S::RPW], gzs[S::RPW], uf[S::RPW], vf[S::RPW], gz_old[S::RPW];
  double rec[S::NREC];
};

template <class S, class PT, class RT>
__device__ __forceinline__ TileIn<S> tile_fetch(const Grid &g, RT mrows, const PT &P, int C0, int R0) {
  const int64_t p = g.pitch;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int q = threadIdx.x % kWave;
  const int i = wrap(C0 - HXL + q, g.nlon);
  const int ic = C0 + q - HXL;
  TileIn<S> in;
#pragma unroll
  for (int h = 0; h < S::RPW; h++) {
    const int k = w + h * S::NW, r = R0 - 2 + k;
    in.gz[h] = in.gzs[h] = in.uf[h] = in.vf[h] = in.gz_old[h] = 0.0;
    if (k >= 2 && k < S::TY + 2 && q >= HXL && q < HXL + TX && ic < g.nlon && r < g.nlat)
      in.gz_old[h] = P.gz_old[r * p + ic];
    if (r >= 0 && r < g.nlat) {
      in.gz[h] = P.gz_star[r * p + i];
      in.gzs[h] = P.gzs[r * p + i];
      in.uf[h] = P.u_f[r * p + i];
    }
    if (r >= 0 && r < g.nhalf) in.vf[h] = P.v_f[r * p + i];
  }
#pragma unroll
  for (int j = 0; j < S::NREC; j++) {
    const int t = threadIdx.x + j * S::NT;
    in.rec[j] = 0.0;
    if (t < S::SH_MESH) {
      const int rr = R0 - 2 + t / kRowRec;
      if (rr >= 0 && rr < g.nlat) in.rec[j] = mrows[(int64_t)(rr + 1) * kRowRec + t % kRowRec];
    }
  }
  return in;
}

// Pole sums of the star state into the tile's pole slots, in the tiles that
// read them (a uniform branch per workgroup).  Runs before the tile's loads
// are fetched: it uses the tile's LDS as scratch.
template <class S, class MT, class PT>
__device__ __forceinline__ void tile_poles(const Grid &g, const MT &Mg, double grav, double rgrav, const PT &P, int R0,
                                           double *sh) {
  double *pole = sh + S::SH_TILES + S::SH_MESH;
  if (R0 <= 2) pole_sums<S::NT>(g, Mg, grav, rgrav, P, 0, sh, pole);
  if (R0 + S::TY + 3 >= g.nlat - 1) pole_sums<S::NT>(g, Mg, grav, rgrav, P, 1, sh, pole);
}

// The stage's intermediates of the star state that the adjoint reads (the
// fields operators_prepare leaves, on the rows it writes them): written by
// the taping tile for its own output points.
// One tile of the forward tendencies and the unfiltered gz update: output
// columns C0 .. C0+TX-1, rows R0 .. R0+TY-1, LDS scratch sh[S::SH_ALL].  The
// tile's pole slots hold its pole sums (tile_poles) and io its loads
// (tile_fetch) on entry.  With TAPE the tile also writes the intermediates
// of its output points to *tape.
// the six upwind derivatives of one edge point into a tape
__device__ __forceinline__ void tape_upwind(double *const *dst, int64_t o, const UpwD &d) {
  dst[0][o] = d.dt;
  dst[1][o] = d.dn;
#pragma unroll
  for (int j = 0; j < 4; j++) dst[2 + j][o] = d.df[j];
}

template <class S, bool TAPE = false, class MT, class PT>
__device__ __forceinline__ void tend_tile(const Grid &g, const MT &Mg, double grav, double rgrav,
                                          double dt, int stokes, int pv_scheme, int order, double wgt, int coriolis,
                                          const PT &P, int C0, int R0, double *sh, const TileIn<S> &io,
                                          const TapeOut *tape = nullptr) {
  double(*sm)[LC] = (double(*)[LC])(sh + 0 * S::NE);
  double(*suf)[LC] = (double(*)[LC])(sh + 1 * S::NE);
  double(*svf)[LC] = (double(*)[LC])(sh + 2 * S::NE);
  double(*smfx)[LC] = (double(*)[LC])(sh + 3 * S::NE);
  double(*smfy)[LC] = (double(*)[LC])(sh + 4 * S::NE);
  double(*spv)[LC] = (double(*)[LC])(sh + 5 * S::NE);
  double(*ske)[LC] = (double(*)[LC])(sh + 6 * S::NE);
  double(*spl)[LC] = suf;  // pv_lon replaces u_f element by element (phase D)
  double(*spt)[LC] = svf;  // pv_lat replaces v_f element by element (phase D)
  double *sme = sh + S::SH_TILES;
  const double *pole = sh + S::SH_TILES + S::SH_MESH;

  const int64_t p = g.pitch;
  const int nlat = g.nlat, nhalf = g.nhalf;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);   // wavefront: tile rows w, w + NW, ..
  const int q = threadIdx.x % kWave;                                   // tile column of this lane
  const int i = wrap(C0 - HXL + q, g.nlon);                            // global column
  const int ic = C0 + q - HXL;                                         // output column of this lane (phase E)
  // tile row k of the wavefront's h-th row, its global row (wave-uniform)
  // and whether the lane holds an output point there
#define TILE_ROW(h)                                   \
  const int k = w + (h) * S::NW;                      \
  const int r = R0 - 2 + k;                           \
  const bool out_pt = k >= 2 && k < S::TY + 2 && q >= HXL && q < HXL + TX && ic < g.nlon && r < nlat;

  // A: m, u_f, v_f of the wavefront's tile rows and the tile's metric rows
#pragma unroll
  for (int h = 0; h < S::RPW; h++) {
    TILE_ROW(h)
    sm[k][q] = (r >= 0 && r < nlat) ? div_rn(io.gz[h] - io.gzs[h], grav, rgrav) : 0.0;   // m_of
    suf[k][q] = io.uf[h];
    svf[k][q] = io.vf[h];
    if (TAPE && out_pt && tape->u_f) {
      tape->u_f[r * p + ic] = io.uf[h];
      if (r < nhalf) tape->v_f[r * p + ic] = io.vf[h];
    }
  }
#pragma unroll
  for (int j = 0; j < S::NREC; j++) {
    const int t = threadIdx.x + j * S::NT;
    if (t < S::SH_MESH) sme[t] = io.rec[j];
  }
  const MeshTile M{sme, R0 - 2};
  __syncthreads();

  // B: mass fluxes (rows 1 .. TY+2), pv (rows 0 .. TY+3), ke (rows 2 .. TY+2).
  // Tile rows that need all three away from the poles form them in one block,
  // so that the compiler interleaves the three dependency chains; the lanes
  // that block adds (mass fluxes on lane 1, ke on lanes 1, 2 and 62) write
  // tile elements that no later phase reads.
#pragma unroll
  for (int h = 0; h < S::RPW; h++) {
  TILE_ROW(h)
  if (!TAPE && k >= 2 && k < S::TY + 3 && r >= 1 && r <= nlat - 3) {
    if (q >= 1 && q < 63) {
      const double m0 = sm[k][q], m1 = sm[k][q + 1], m2 = sm[k + 1][q], m3 = sm[k + 1][q + 1];
      const double u0 = suf[k][q], u1 = suf[k + 1][q], uw = suf[k][q - 1];
      const double v0 = svf[k][q], v1 = svf[k][q + 1], vs = svf[k - 1][q];
      smfx[k][q] = (m0 + m1) * 0.5 * u0;
      smfy[k][q] = (m0 + m2) * 0.5 * v0;
      spv[k][q] = pv_of(vor_of(u0, u1, v1, v0, M, r), M, r, m_vtx_of(m0, m1, m2, m3, M, r));
      ske[k][q] = ke_of(uw, u0, vs, v0, M, r);
    }
  } else {
  if (k >= 1 && k < S::TY + 3 && q >= 2 && q < 63) {
    smfx[k][q] = (r >= 1 && r <= nlat - 2) ? (sm[k][q] + sm[k][q + 1]) * 0.5 * suf[k][q] : 0.0;
    smfy[k][q] = (r >= 0 && r < nhalf) ? (sm[k][q] + sm[k + 1][q]) * 0.5 * svf[k][q] : 0.0;
  }
  if (k < S::TY + 4 && q >= 1 && q < 63) {
    double v = 0.0;
    if (r >= 0 && r < nhalf) {
      double vor;
      if (stokes && r == 0) vor = pole[4];
      else if (stokes && r == nhalf - 1) vor = pole[5];
      else vor = vor_of(suf[k][q], suf[k + 1][q], svf[k][q + 1], svf[k][q], M, r);
      const double mv = m_vtx_of(sm[k][q], sm[k][q + 1], sm[k + 1][q], sm[k + 1][q + 1], M, r);
      v = pv_of(vor, M, r, mv);
      if (TAPE && out_pt) {
        tape->m_vtx[r * p + ic] = mv;
        tape->pv[r * p + ic] = v;
        tape->mfy[r * p + ic] = smfy[k][q];
        tape->m_lat[r * p + ic] = (sm[k][q] + sm[k + 1][q]) * 0.5;
      }
    } else if (r >= -2 && r <= nhalf + 1) {
      int rr = r, ii = i;
      mirror_half(rr, ii, g);
      v = pv_global(g, Mg, grav, rgrav, stokes, pole, P, rr, ii);
    }
    spv[k][q] = v;
  }
  if (k >= 2 && k < S::TY + 3 && q >= 3 && q < 62) {
    double v = 0.0;
    if (r == 0) v = pole[0];
    else if (r == nlat - 1) v = pole[1];
    else if (r > 0 && r < nlat - 1) v = ke_of(suf[k][q - 1], suf[k][q], svf[k - 1][q], svf[k][q], M, r);
    ske[k][q] = v;
  }
  }
  }
  __syncthreads();

  // D: tangential winds and pv on the edges.  v_lon (U rows 2 .. TY+2) and
  // u_lat (V rows 1 .. TY+1) are each read only by the lane that forms them,
  // so they stay in registers; pv_lon replaces u_f and pv_lat replaces v_f
  // element by element (each lane reads only its own u_f / v_f here, in each
  // of its rows).
#pragma unroll
  for (int h = 0; h < S::RPW; h++) {
  TILE_ROW(h)
  double pl = 0.0, pt = 0.0;
  if (!TAPE && k >= 2 && k < S::TY + 2 && r >= 1 && r <= nlat - 2) {
    // both edges in one block (interleaved chains); lane 2 forms a pv_lat
    // that is replaced by the 0 the edge loop below leaves there
    if (q >= 2 && q < 62) {
      const double tl = tangent_sum(M(GMC_FULL_TANGENT_WGT1, r), M(GMC_FULL_TANGENT_WGT2, r), smfy[k - 1][q],
                                    smfy[k - 1][q + 1], smfy[k][q], smfy[k][q + 1]);
      const double tu = tangent_sum(M(GMC_HALF_TANGENT_WGT1, r), M(GMC_HALF_TANGENT_WGT2, r), smfx[k][q - 1],
                                    smfx[k][q], smfx[k + 1][q - 1], smfx[k + 1][q]);
      const double vl = tl / ((sm[k][q] + sm[k][q + 1]) * 0.5);
      const double ul = tu / ((sm[k][q] + sm[k + 1][q]) * 0.5);
      const double el = pv_edge(pv_scheme, order, wgt, vl, suf[k][q], 0, dt, M(GMC_LE_LON, r), spv[k - 2][q],
                                spv[k - 1][q], spv[k][q], spv[k + 1][q]);
      const double et = pv_edge(pv_scheme, order, wgt, ul, svf[k][q], 1, dt, M(GMC_LE_LAT, r), spv[k][q - 2],
                                spv[k][q - 1], spv[k][q], spv[k][q + 1]);
      pl = el;
      pt = q >= 3 ? et : 0.0;
    }
  } else {
  if (k >= 2 && k < S::TY + 3 && q >= 2 && q < 62 && r >= 1 && r <= nlat - 2) {
    const double tt = tangent_sum(M(GMC_FULL_TANGENT_WGT1, r), M(GMC_FULL_TANGENT_WGT2, r), smfy[k - 1][q],
                                  smfy[k - 1][q + 1], smfy[k][q], smfy[k][q + 1]);
    const double vl = tt / ((sm[k][q] + sm[k][q + 1]) * 0.5);
    const double a1 = spv[k - 2][q], a2 = spv[k - 1][q], a3 = spv[k][q], a4 = spv[k + 1][q];
    pl = pv_edge(pv_scheme, order, wgt, vl, suf[k][q], 0, dt, M(GMC_LE_LON, r), a1, a2, a3, a4);
    if (TAPE && out_pt) {
      tape->m_lon[r * p + ic] = (sm[k][q] + sm[k][q + 1]) * 0.5;
      tape->mfx[r * p + ic] = smfx[k][q];
      if (tape->mfy_lon) tape->mfy_lon[r * p + ic] = tt;
      tape->v_lon[r * p + ic] = vl;
      tape->pv_lon[r * p + ic] = pl;
      if (tape->dlon[0]) tape_upwind(tape->dlon, r * p + ic,
                                     upwind_deriv(pv_scheme, order, wgt, vl, suf[k][q], 0, a1, a2, a3, a4));
    }
  }
  if (k >= 1 && k < S::TY + 2 && q >= 3 && q < 62 && r >= 0 && r < nhalf) {
    const double tt = tangent_sum(M(GMC_HALF_TANGENT_WGT1, r), M(GMC_HALF_TANGENT_WGT2, r), smfx[k][q - 1],
                                  smfx[k][q], smfx[k + 1][q - 1], smfx[k + 1][q]);
    const double ul = tt / ((sm[k][q] + sm[k + 1][q]) * 0.5);
    const double a1 = spv[k][q - 2], a2 = spv[k][q - 1], a3 = spv[k][q], a4 = spv[k][q + 1];
    pt = pv_edge(pv_scheme, order, wgt, ul, svf[k][q], 1, dt, M(GMC_LE_LAT, r), a1, a2, a3, a4);
    if (TAPE && out_pt) {
      if (tape->mfx_lat) tape->mfx_lat[r * p + ic] = tt;
      tape->u_lat[r * p + ic] = ul;
      tape->pv_lat[r * p + ic] = pt;
      if (tape->dlat[0]) tape_upwind(tape->dlat, r * p + ic,
                                     upwind_deriv(pv_scheme, order, wgt, ul, svf[k][q], 1, a1, a2, a3, a4));
    }
  }
  }
  spl[k][q] = pl;
  spt[k][q] = pt;
  }
  __syncthreads();

  // E: du, dv (forward pass) and gz_new = gz_old + dt dgz on the output rows;
  // rows away from the poles in one block
#pragma unroll
  for (int h = 0; h < S::RPW; h++) {
  TILE_ROW(h)
  const double gz_old = io.gz_old[h];
  if (r >= 1 && r <= nlat - 2) {
    if (out_pt) {
      const int64_t o = r * p + ic;
      const double qhv = coriolis == 1
          ? qhv_of(M, r, spl[k][q], smfy[k - 1][q], spt[k - 1][q], smfy[k - 1][q + 1], spt[k - 1][q + 1],
                   smfy[k][q], spt[k][q], smfy[k][q + 1], spt[k][q + 1])
          : tangent_sum(M(GMC_FULL_TANGENT_WGT1, r), M(GMC_FULL_TANGENT_WGT2, r), smfy[k - 1][q],
                        smfy[k - 1][q + 1], smfy[k][q], smfy[k][q + 1]) * spl[k][q];
      const double dku = M.div(ske[k][q + 1] - ske[k][q], GMC_DE_LON, r);
      const double qhu = coriolis == 1
          ? qhu_of(M, r, spt[k][q], smfx[k][q - 1], spl[k][q - 1], smfx[k][q], spl[k][q],
                   smfx[k + 1][q - 1], spl[k + 1][q - 1], smfx[k + 1][q], spl[k + 1][q])
          : tangent_sum(M(GMC_HALF_TANGENT_WGT1, r), M(GMC_HALF_TANGENT_WGT2, r), smfx[k][q - 1], smfx[k][q],
                        smfx[k + 1][q - 1], smfx[k + 1][q]) * spt[k][q];
      const double dkv = M.div(ske[k + 1][q] - ske[k][q], GMC_DE_LAT, r);
      const double dgz = -(dmfdlon_of(smfx[k][q], smfx[k][q - 1], M, r) +
                           dmfdlat_of(smfy[k][q], smfy[k - 1][q], M, r)) * grav;
      P.du[o] = qhv - dku;
      P.dv[o] = -qhu - dkv;
      P.gz_new[o] = gz_old + dt * dgz;
    }
    continue;
  }
  // the pole rows: v on the southern pole edge row, gz from the pole sums
  if (out_pt) {
    const int64_t o = r * p + ic;
    if (r < nhalf) {
      const double qhu = coriolis == 1
          ? qhu_of(M, r, spt[k][q], smfx[k][q - 1], spl[k][q - 1], smfx[k][q], spl[k][q],
                   smfx[k + 1][q - 1], spl[k + 1][q - 1], smfx[k + 1][q], spl[k + 1][q])
          : tangent_sum(M(GMC_HALF_TANGENT_WGT1, r), M(GMC_HALF_TANGENT_WGT2, r), smfx[k][q - 1], smfx[k][q],
                        smfx[k + 1][q - 1], smfx[k + 1][q]) * spt[k][q];
      const double dk = M.div(ske[k + 1][q] - ske[k][q], GMC_DE_LAT, r);
      P.dv[o] = -qhu - dk;
    }
    const double dmfdlon = 0.0, dmfdlat = r == 0 ? pole[2] : pole[3];
    const double dgz = -(dmfdlon + dmfdlat) * grav;
    P.gz_new[o] = gz_old + dt * dgz;
  }
  }
#undef TILE_ROW
}

// Forward tendencies and the unfiltered gz update on a 58 x TY_MAIN tile.
// Wavefront w owns tile rows k = w, w + NW, .. (global row R0 - 2 + k) in
// every phase and lane q the column C0 - 3 + q, so metric terms are
// wave-uniform (broadcast LDS reads of the tile's metric rows) and each phase
// is one pass over each of its rows.  Blocks are dealt to the XCDs
// round-robin; each XCD gets a contiguous run of tiles (row-major), so
// vertically adjacent tiles share that XCD's L2.  The XCD holding the north
// pole walks its run backwards so that the polar tiles, which also form the
// pole sums, start first.
constexpr int kTendThreads = Tile<TY_MAIN, RPW_MAIN>::NT;
template <bool TAPE>
__global__ void __launch_bounds__(kTendThreads, 8 / RPW_MAIN) k_tend(Grid g, Mesh Mg, const double *__restrict__ mrows,
                                                             double grav, double rgrav, double dt, int stokes,
                                                             int pv_scheme, int order, double wgt, int coriolis,
                                                             StagePtrs P, TapeOut tape) {
  using S = Tile<TY_MAIN, RPW_MAIN>;
  __shared__ double sh[S::SH_ALL];
  // tile of this block
  const int nx = (g.nlon + TX - 1) / TX, ny = (g.nlat + S::TY - 1) / S::TY, T = nx * ny;
  const int b = blockIdx.x, x = b % NXCD, s = b / NXCD;
  int start = 0;
  for (int y = 0; y < x; y++) start += (T - y + NXCD - 1) / NXCD;
  const int cnt = (T - x + NXCD - 1) / NXCD;
  const int t = start + (x == NXCD - 1 ? cnt - 1 - s : s);
  const int C0 = (t % nx) * TX, R0 = (t / nx) * S::TY;

  tile_poles<S>(g, Mg, grav, rgrav, P, R0, sh);
  const TileIn<S> io = tile_fetch<S>(g, mrows, P, C0, R0);
  tend_tile<S, TAPE>(g, Mg, grav, rgrav, dt, stokes, pv_scheme, order, wgt, coriolis, P, C0, R0, sh, io, &tape);
}

// ---------------------------------------------------------------------------
// zonal filters over (row, column chunk) work items
struct Item {
  int32_t row, c0, c1, kind;  // kind: 0 gz (cell row), 1 u (U row), 2 v (V row)
  int32_t member;             // ensemble member (0 for a single context)
  int32_t n, n2;              // taps of the row's filters (copies of the filter tables): the big filter
                              // of the item's row kind, and small_filter1 for gz items
  int32_t pad;
};

// column i - h + t of a row, t in [0, nlon + 2h), h <= nlon
__device__ __forceinline__ int wrap2(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

// big filter gz_new -> gz_f and small filter gz_new -> gz
// (time_schemes_mod.F90:184-191, filter_mod.F90:25-50).  Weights and the
// row window are staged in LDS: xs = [weights big | weights small | window].
template <class FT, class CP, class WP>
__device__ __forceinline__ void gzf_item(const Grid &g, const FT &big, const FT &small, const Item &it, CP gz_new,
                                         WP gz_f, WP gz, double *xs) {
  const int r = it.row;
  const int64_t o = r * g.pitch;
  const int nb = it.n, ns = it.n2;
  const int hb = nb > 1 ? (nb - 1) / 2 : 0, hs = ns > 1 ? (ns - 1) / 2 : 0, hw = hb > hs ? hb : hs;
  double *wb = xs, *ws = xs + big.max_taps, *x0 = xs + big.max_taps + small.max_taps;
  const int len = it.c1 - it.c0 + 2 * hw;
  for (int k = threadIdx.x; k < nb; k += NF) wb[k] = big.wgt_lon[(int64_t)r * big.max_taps + k];
  for (int k = threadIdx.x; k < ns; k += NF) ws[k] = small.wgt_lon[(int64_t)r * small.max_taps + k];
  for (int t = threadIdx.x; t < len; t += 4 * NF) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (t + u * NF < len) v[u] = gz_new[o + wrap2(it.c0 - hw + t + u * NF, g.nlon)];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (t + u * NF < len) x0[t + u * NF] = v[u];
  }
  __syncthreads();
  for (int c = it.c0 + threadIdx.x; c < it.c1; c += NF) {
    const double *x = x0 + (c - it.c0);      // x[hw] is column c
    gz_f[o + c] = nb > 1 ? filter_point(wb, nb, x + hw - hb) : x[hw];
    gz[o + c] = ns > 1 ? filter_point(ws, ns, x + hw - hs) : x[hw];
  }
}

__global__ void __launch_bounds__(NF) k_gzf(Grid g, FilterDev big, FilterDev small, const Item *__restrict__ items,
                                            const double *__restrict__ gz_new, double *gz_f, double *gz) {
  extern __shared__ double xs[];
  gzf_item(g, big, small, items[xcd_item(blockIdx.x, gridDim.x)], gz_new, gz_f, gz, xs);
}

struct UVPtrs {
  const double *u_old, *v_old, *du, *dv, *gz_f;
  double *u, *v, *u_f, *v_f;
};

// backward pass + u/v update + big filters
// (gmcore_mod.F90:552-571, time_schemes_mod.F90:198-239).  xs = [weights |
// window of the updated row].
template <class MT, class FT, class PT>
__device__ __forceinline__ void uvf_item(const Grid &g, const MT &M, double dt, const FT &big,
                                         const Item &it, const PT &P, double *xs) {
  const int r = it.row;
  const int64_t p = g.pitch, o = r * p;
  const int nl = g.nlon;
  double *w = xs, *x0 = xs + big.max_taps;
  if (it.kind == 1 && (r == 0 || r == g.nlat - 1)) {   // pole U rows: not updated, copied to u_f
    for (int c = it.c0 + threadIdx.x; c < it.c1; c += NF) P.u_f[o + c] = P.u[o + c];
    return;
  }
  const int n = it.n;
  const int h = n > 1 ? (n - 1) / 2 : 0;
  const int len = it.c1 - it.c0 + 2 * h;
  const auto wg = (it.kind == 1 ? big.wgt_lon : big.wgt_lat) + (int64_t)r * big.max_taps;
  for (int k = threadIdx.x; k < n; k += NF) w[k] = wg[k];
  if (it.kind == 1) {
    const double de = M(GMC_DE_LON, r), rde = 1.0 / de;
    for (int t = threadIdx.x; t < len; t += 4 * NF) {
      double uo[4], du[4], ge[4], gw[4];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (t + u * NF < len) {
          const int i = wrap2(it.c0 - h + t + u * NF, nl);
          uo[u] = P.u_old[o + i];
          du[u] = P.du[o + i];
          ge[u] = P.gz_f[o + wrap(i + 1, nl)];
          gw[u] = P.gz_f[o + i];
        }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (t + u * NF < len) x0[t + u * NF] = uo[u] + dt * (du[u] - div_rn(ge[u] - gw[u], de, rde));
    }
  } else {
    const int pole = r == 0 ? 1 : (r == g.nhalf - 1 ? -1 : 0);   // neighbour row of the smoothing
    const double de0 = M(GMC_DE_LAT, r), de1 = pole ? M(GMC_DE_LAT, r + pole) : 1.0;
    const double rde0 = 1.0 / de0, rde1 = 1.0 / de1;
    for (int t = threadIdx.x; t < len; t += 2 * NF) {
      double a[2][8];
#pragma unroll
      for (int u = 0; u < 2; u++)
        if (t + u * NF < len) {
          const int64_t q = o + wrap2(it.c0 - h + t + u * NF, nl);
          a[u][0] = P.v_old[q]; a[u][1] = P.dv[q]; a[u][2] = P.gz_f[q + p]; a[u][3] = P.gz_f[q];
          if (pole) {
            const int64_t q1 = q + pole * p;
            a[u][4] = P.v_old[q1]; a[u][5] = P.dv[q1]; a[u][6] = P.gz_f[q1 + p]; a[u][7] = P.gz_f[q1];
          }
        }
#pragma unroll
      for (int u = 0; u < 2; u++)
        if (t + u * NF < len) {
          double v = a[u][0] + dt * (a[u][1] - div_rn(a[u][2] - a[u][3], de0, rde0));
          if (pole) v = pole_smooth(v, a[u][4] + dt * (a[u][5] - div_rn(a[u][6] - a[u][7], de1, rde1)));
          x0[t + u * NF] = v;
        }
    }
  }
  __syncthreads();
  const auto out = it.kind == 1 ? P.u : P.v;
  const auto out_f = it.kind == 1 ? P.u_f : P.v_f;
  for (int c = it.c0 + threadIdx.x; c < it.c1; c += NF) {
    const double *x = x0 + (c - it.c0);
    out[o + c] = x[h];
    out_f[o + c] = n > 1 ? filter_point(w, n, x) : x[h];
  }
}

__global__ void __launch_bounds__(NF) k_uvf(Grid g, Mesh M, double dt, FilterDev big, const Item *__restrict__ items,
                                            UVPtrs P) {
  extern __shared__ double xs[];
  uvf_item(g, M, dt, big, items[xcd_item(blockIdx.x, gridDim.x)], P, xs);
}


// ---------------------------------------------------------------------------
// Batched ensembles: the same stage kernels over many contexts in one launch.
// A member's descriptor for one stage holds its geometry, tables, options,
// stage dt and field pointers; tiles and filter items of all members are
// numbered consecutively and each workgroup finds its member by the tile
// offset (k_tend) or the item's member index (filters).
struct EnsMember {
  Grid g;
  Mesh M;
  const double *mrows;
  FilterDev big, small1;
  double grav, rgrav, dt, wgt;   // rgrav = 1 / grav correctly rounded
  int stokes, pv_scheme, order, coriolis;
  StagePtrs sp;
  UVPtrs up;
  double *gz_f_out;    // gz_f of the new state (big filter output)
  double *gz_out;      // gz of the new state (small filter output)
  int tile0, nx;       // first tile of the member, tiles per tile row
};

// Field pointers read from a descriptor in memory are generic (FLAT) to the
// compiler, and FLAT accesses also count against the LDS wait counter.  The
// ensemble kernels therefore hand the shared tile and item code pointers
// typed in the global address space, so that it emits global loads and
// stores as it does for pointers passed as kernel arguments.
template <class T> using gptr = __attribute__((address_space(1))) T *;
template <class T> __device__ __forceinline__ gptr<T> glb(T *p) { return (gptr<T>)p; }
struct MeshG {
  gptr<const double> m;
  int nlat;
  __device__ __forceinline__ double operator()(int f, int r) const { return m[f * nlat + r]; }
  __device__ __forceinline__ double div(double x, int f, int r) const { return x / (*this)(f, r); }
};
__device__ __forceinline__ MeshG glb(const Mesh &M) { return MeshG{glb(M.m), M.nlat}; }
struct FilterDevG {
  int max_taps;
  gptr<const int32_t> ngrid_lon, ngrid_lat;
  gptr<const double> wgt_lon, wgt_lat;
};
struct StagePtrsG {
  gptr<const double> gz_star, gzs, u_f, v_f, gz_old;
  gptr<double> gz_new, du, dv;
};
struct UVPtrsG {
  gptr<const double> u_old, v_old, du, dv, gz_f;
  gptr<double> u, v, u_f, v_f;
};
__device__ __forceinline__ FilterDevG glb(const FilterDev &f) {
  return FilterDevG{f.max_taps, glb(f.ngrid_lon), glb(f.ngrid_lat), glb(f.wgt_lon), glb(f.wgt_lat)};
}
__device__ __forceinline__ StagePtrsG glb(const StagePtrs &p) {
  return StagePtrsG{glb(p.gz_star), glb(p.gzs), glb(p.u_f), glb(p.v_f), glb(p.gz_old), glb(p.gz_new), glb(p.du), glb(p.dv)};
}
__device__ __forceinline__ UVPtrsG glb(const UVPtrs &p) {
  return UVPtrsG{glb(p.u_old), glb(p.v_old), glb(p.du), glb(p.dv), glb(p.gz_f), glb(p.u), glb(p.v), glb(p.u_f), glb(p.v_f)};
}

// The tiles of all members, numbered member by member and row-major within a
// member, are cut into NXCD contiguous runs, one per XCD (workgroup b runs on
// XCD b % NXCD), so vertically adjacent tiles share that XCD's L2.  The
// workgroups of one XCD stride over its run: gridDim.x may be smaller than
// the tile count (a persistent launch sized by the tuner).
template <class S>
__global__ void __launch_bounds__(S::NT, 8) k_tend_ens(const EnsMember *__restrict__ mem, int nmem, int total) {
  __shared__ double sh[S::SH_ALL];
  const int G = gridDim.x, x = blockIdx.x % NXCD;
  int start = 0;
  for (int y = 0; y < x; y++) start += (total - y + NXCD - 1) / NXCD;
  const int cnt = (total - x + NXCD - 1) / NXCD;    // tiles of this XCD's run
  const int per = (G - x + NXCD - 1) / NXCD;        // workgroups on this XCD
  for (int s = blockIdx.x / NXCD; s < cnt; s += per) {
    const int t = start + s;
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (mem[mid].tile0 <= t) lo = mid; else hi = mid - 1;
    }
    const EnsMember d = mem[lo];   // a private copy: no aliasing with the stores of the tile
    const int lt = t - d.tile0;
    const int C0 = (lt % d.nx) * TX, R0 = (lt / d.nx) * S::TY;
    const StagePtrsG sp = glb(d.sp);
    tile_poles<S>(d.g, glb(d.M), d.grav, d.rgrav, sp, R0, sh);
    const TileIn<S> io = tile_fetch<S>(d.g, glb(d.mrows), sp, C0, R0);
    tend_tile<S>(d.g, glb(d.M), d.grav, d.rgrav, d.dt, d.stokes, d.pv_scheme, d.order, d.wgt, d.coriolis, sp, C0,
                 R0, sh, io);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(NF) k_gzf_ens(const EnsMember *__restrict__ mem, const Item *__restrict__ items,
                                                int nitems) {
  extern __shared__ double xs[];
  for (int b = blockIdx.x; b < nitems; b += gridDim.x) {
    const Item it = items[b];
    const EnsMember d = mem[it.member];
    gzf_item(d.g, glb(d.big), glb(d.small1), it, glb(d.sp.gz_new), glb(d.gz_f_out), glb(d.gz_out), xs);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(NF) k_uvf_ens(const EnsMember *__restrict__ mem, const Item *__restrict__ items,
                                                int nitems) {
  extern __shared__ double xs[];
  for (int b = blockIdx.x; b < nitems; b += gridDim.x) {
    const Item it = items[b];
    const EnsMember *d = mem + it.member;
    const Grid g = d->g;
    const MeshG M = glb(d->M);
    const FilterDevG big = glb(d->big);
    const UVPtrsG up = glb(d->up);
    uvf_item(g, M, d->dt, big, it, up, xs);
    __syncthreads();
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Work lists: a row whose filter has n taps is cut into chunks of about
// kChunkWork / n columns (a multiple of 64, at least 64) so that each
// workgroup does similar work: the taps of a chunk's outputs are LDS reads,
// and a polar row of several hundred taps over 512 columns kept one CU busy
// for the whole launch.  Rows of short filters go into chunks of at most
// kNarrowChunk columns, so that a workgroup's window is one round of loads.
static constexpr int kChunkWork = GMC_CHUNK_WORK;
static constexpr int kNarrowChunk = GMC_NARROW_CHUNK;

// One row of a filter work list: `work` taps per output sets the chunk
// width; n (and n2) are the taps of the row's filters, the window of a chunk
// is its columns and (n - 1) / 2 on either side (lds: the largest window).
// Returns the number of chunks the row was cut into.
static int add_row(std::vector<Item> &v, int row, int kind, int work, int nlon, size_t &lds, int n, int n2 = 0) {
  int w = work <= 16 ? std::min(nlon, kNarrowChunk)
                     : std::min(nlon, std::max(64, (kChunkWork / work + 63) / 64 * 64));
  const int parts = (nlon + w - 1) / w;
  const int h = n > 1 ? (n - 1) / 2 : 0;
  for (int k = 0; k < parts; k++) {
    const int c0 = (int)((int64_t)nlon * k / parts), c1 = (int)((int64_t)nlon * (k + 1) / parts);
    v.push_back(Item{row, c0, c1, kind, 0, n, n2, 0});
    lds = std::max(lds, sizeof(double) * (size_t)(c1 - c0 + 2 * h));
  }
  return parts;
}

struct FusedPlan {
  Item *d_items = nullptr;
  int n_gz = 0, n_uv = 0;
  size_t lds_gz = 0, lds_uv = 0;
  int rows_split = 0, max_parts = 1, widest = 0;   // reported by gmc_filter_plan
};

static FusedPlan *plan_of(gmc_ctx *c) { return (FusedPlan *)c->fused_plan; }

// Order a work list for a launch of one workgroup per item whose workgroup b
// takes item xcd_item(b): each XCD works through one contiguous run of the
// list.  The wide items (more than twice the median work: the polar rows'
// long filters) are dealt round-robin to the XCDs and put first in each run,
// so they start first everywhere; the rest follow in their list order, so
// each XCD sweeps a band of neighbouring rows that share its L2.
template <class W>
static std::vector<Item> xcd_deal(const std::vector<Item> &v, W work) {
  const int T = (int)v.size();
  std::vector<int64_t> wk(T);
  for (int k = 0; k < T; k++) wk[k] = work(v[k]);
  std::vector<int64_t> sorted = wk;
  std::nth_element(sorted.begin(), sorted.begin() + T / 2, sorted.end());
  const int64_t med = T ? sorted[T / 2] : 0;
  std::vector<int> wide, narrow;
  for (int k = 0; k < T; k++) (wk[k] > 2 * med ? wide : narrow).push_back(k);
  std::stable_sort(wide.begin(), wide.end(), [&](int a, int b) { return wk[a] > wk[b]; });
  std::vector<std::vector<int>> run(NXCD);
  std::vector<int> size(NXCD);
  for (int x = 0; x < NXCD; x++) size[x] = T / NXCD + (x < T % NXCD ? 1 : 0);   // run lengths of xcd_item
  int x = 0;
  for (int k : wide) {   // round-robin over the XCDs that still have room
    while ((int)run[x].size() >= size[x]) x = (x + 1) % NXCD;
    run[x].push_back(k);
    x = (x + 1) % NXCD;
  }
  size_t next = 0;
  std::vector<Item> out;
  out.reserve(T);
  for (int y = 0; y < NXCD; y++) {
    while ((int)run[y].size() < size[y]) run[y].push_back(narrow[next++]);
    for (int k : run[y]) out.push_back(v[k]);
  }
  return out;
}

static int build_plan(gmc_ctx *c) {
  const Grid &g = c->grid;
  std::vector<int32_t> nb_lon(g.nlat), nb_lat(g.nhalf), ns_lon(g.nlat);
  int rc = check_hip(c, hipMemcpy(nb_lon.data(), c->big.ngrid_lon, sizeof(int32_t) * g.nlat, hipMemcpyDeviceToHost),
                     "plan");
  if (!rc) rc = check_hip(c, hipMemcpy(nb_lat.data(), c->big.ngrid_lat, sizeof(int32_t) * g.nhalf,
                                       hipMemcpyDeviceToHost), "plan");
  if (!rc) rc = check_hip(c, hipMemcpy(ns_lon.data(), c->small1.ngrid_lon, sizeof(int32_t) * g.nlat,
                                       hipMemcpyDeviceToHost), "plan");
  if (rc) return rc;
  for (int r = 0; r < g.nlat; r++)
    if (ns_lon[r] > g.nlon) return set_error(c, GMC_EINVAL, "filter wider than a latitude circle");
  for (int r = 0; r < g.nlat; r++)
    if (nb_lon[r] > g.nlon) return set_error(c, GMC_EINVAL, "filter wider than a latitude circle");
  for (int r = 0; r < g.nhalf; r++)
    if (nb_lat[r] > g.nlon) return set_error(c, GMC_EINVAL, "filter wider than a latitude circle");
  if (g.nlon > Tile<TY_MAIN, RPW_MAIN>::kMaxPoleLon) return set_error(c, GMC_EINVAL, "nlon too large for the fused stage kernels");
  std::vector<Item> items;
  FusedPlan *P = new FusedPlan();

  {
    std::vector<Item> gz, uv;
    auto count = [&](int parts, int n) {
      P->rows_split += parts > 1;
      P->max_parts = std::max(P->max_parts, parts);
      P->widest = std::max(P->widest, n);
    };
    // gz items: the window is that of the wider of the two filters
    for (int r = 0; r < g.nlat; r++)
      count(add_row(gz, r, 0, nb_lon[r] + std::max(1, nb_lon[r] / 2), g.nlon, P->lds_gz,
                    std::max(nb_lon[r], ns_lon[r])), nb_lon[r]);
    for (auto &it : gz) { it.n = nb_lon[it.row]; it.n2 = ns_lon[it.row]; }
    // u row r next to v rows r - 1 and r: the three read gz_f row r
    for (int r = 0; r < g.nlat; r++) {
      count(add_row(uv, r, 1, nb_lon[r], g.nlon, P->lds_uv, nb_lon[r]), nb_lon[r]);
      if (r < g.nhalf) count(add_row(uv, r, 2, nb_lat[r], g.nlon, P->lds_uv, nb_lat[r]), nb_lat[r]);
    }
    auto work = [&](const Item &it) {
      const int n = it.kind == 2 ? nb_lat[it.row] : nb_lon[it.row];
      return (int64_t)(it.c1 - it.c0) * std::max(n, 1);
    };
    const std::vector<Item> gz_d = xcd_deal(gz, work), uv_d = xcd_deal(uv, work);
    items.insert(items.end(), gz_d.begin(), gz_d.end());
    P->n_gz = (int)items.size();
    items.insert(items.end(), uv_d.begin(), uv_d.end());
    P->n_uv = (int)items.size() - P->n_gz;
  }
  P->lds_gz += sizeof(double) * (c->big.max_taps + c->small1.max_taps);
  P->lds_uv += sizeof(double) * c->big.max_taps;
  rc = check_hip(c, hipMalloc((void **)&P->d_items, sizeof(Item) * items.size()), "plan alloc");
  if (!rc) rc = check_hip(c, hipMemcpy(P->d_items, items.data(), sizeof(Item) * items.size(), hipMemcpyHostToDevice),
                          "plan copy");
  if (rc) { delete P; return rc; }
  c->fused_plan = P;
  return GMC_OK;
}

void fused_release(gmc_ctx *c) {
  for (auto &sg : c->graphs) hip_note(hipGraphExecDestroy(sg.exec), "step graph destroy");
  c->graphs.clear();
  if (FusedPlan *P = plan_of(c)) {
    hip_note(hipFree(P->d_items), "plan free");
    delete P;
    c->fused_plan = nullptr;
  }
}

const char *const kKernelNames[kKernelClasses] = {"k_tend", "k_gzf", "k_uvf"};

// A stage launch.  Under gmc_kernel_times the next two profiling events take
// the kernel's own start and end stamps (hipExtLaunchKernelGGL fills them from
// the dispatch itself, so no event packet sits between them and the kernel);
// every argument is converted to the kernel's parameter type first, since the
// extended launch passes the arguments' bytes as given.
template <class... P, class... A>
static void stage_launch(gmc_ctx *c, void (*k)(P...), dim3 grid, dim3 block, uint32_t lds, A &&...a) {
  static_assert(sizeof...(P) == sizeof...(A), "one argument per kernel parameter");
  if (c->prof_ev && c->prof_n + 2 <= c->prof_cap) {
    hipExtLaunchKernelGGL(k, grid, block, lds, c->stream, c->prof_ev[c->prof_n], c->prof_ev[c->prof_n + 1], 0,
                          static_cast<P>(a)...);
    c->prof_n += 2;
  } else {
    hipLaunchKernelGGL(k, grid, block, lds, c->stream, static_cast<P>(a)...);
  }
}

// one step_forward_backward stage: old -> new with the star state `star`;
// with tape, the star state's intermediates go to its state fields
static void fused_stage(gmc_ctx *c, int o, int s, int n, double dt, bool tape = false,
                        const TapeOut *dst = nullptr) {
  const Grid &g = c->grid;
  const Mesh M{c->d_mesh, g.nlat};
  FusedPlan *PL = plan_of(c);
  double **O = c->state[o], **S = c->state[s], **N = c->state[n];
  double *du = c->tend[o][T_DU], *dv = c->tend[o][T_DV];
  double *gz_new = c->tend[o][T_DGZ];   // the unfiltered gz of the new state
  StagePtrs sp{S[S_GZ], c->gzs, S[S_U_F], S[S_V_F], O[S_GZ], gz_new, du, dv};
  const int tiles = ((g.nlon + TX - 1) / TX) * ((g.nlat + TY_MAIN - 1) / TY_MAIN);
  TapeOut T{S[S_M_LON], S[S_M_LAT], S[S_M_VTX], S[S_MFX_LON], S[S_MFY_LAT], S[S_MFX_LAT], S[S_MFY_LON],
            S[S_U_LAT], S[S_V_LON], S[S_PV], S[S_PV_LON], S[S_PV_LAT], nullptr, nullptr};
  if (dst) T = *dst;
  stage_launch(c, tape ? k_tend<true> : k_tend<false>, dim3(tiles), dim3(Tile<TY_MAIN, RPW_MAIN>::NT), 0, g, M,
               c->d_mrows, c->cfg.g, 1.0 / c->cfg.g, dt, c->cfg.pv_pole_stokes, c->cfg.pv_scheme,
               pv_order(c->cfg), c->cfg.upwind_wgt_pv, c->cfg.coriolis_scheme, sp, T);
  stage_launch(c, k_gzf, dim3(PL->n_gz), dim3(NF), PL->lds_gz, g, c->big, c->small1, PL->d_items, gz_new,
               N[S_GZ_F], N[S_GZ]);
  UVPtrs up{O[S_U_LON], O[S_V_LAT], du, dv, N[S_GZ_F], N[S_U_LON], N[S_V_LAT], N[S_U_F], N[S_V_F]};
  stage_launch(c, k_uvf, dim3(PL->n_uv), dim3(NF), PL->lds_uv, g, M, dt, c->big, PL->d_items + PL->n_gz, up);
}

int fused_prepare(gmc_ctx *c) { return c->fused_plan ? GMC_OK : build_plan(c); }

int fused_plan_info(gmc_ctx *c, int32_t out[4]) {
  const int rc = fused_prepare(c);
  if (rc) return rc;
  const FusedPlan *P = plan_of(c);
  out[0] = P->n_gz + P->n_uv;
  out[1] = P->rows_split;
  out[2] = P->max_parts;
  out[3] = P->widest;
  return GMC_OK;
}

int fused_stage_tape(gmc_ctx *c, int o, int s, int n, double dt, const TapeOut *dst) {
  int rc = fused_prepare(c);
  if (rc) return rc;
  fused_stage(c, o, s, n, dt, true, dst);
  return check_hip(c, hipGetLastError(), "fused stage");
}

int fused_integrate(gmc_ctx *c, int o, int n, double dt) {
  int rc = fused_prepare(c);
  if (rc) return rc;
  const bool keep = c->keep_derived != 0;
  switch (c->cfg.time_scheme) {
  case GMC_TIME_EULER:
    fused_stage(c, o, o, n, dt, keep);
    break;
  case GMC_TIME_WRFRK3:
    fused_stage(c, o, o, n, dt / 3.0, keep);
    fused_stage(c, o, n, 3, dt / 2.0, keep);
    fused_stage(c, o, 3, n, dt, keep);
    break;
  default:
    fused_stage(c, o, o, n, dt / 2.0, keep);
    fused_stage(c, o, n, 3, dt / 2.0, keep);
    fused_stage(c, o, 3, n, dt, keep);
  }
  rc = check_hip(c, hipGetLastError(), "fused stage");
  if (!rc && c->cfg.use_div_damp) rc = ops_div_damp(c, n);
  return rc;
}

void derived_after_step(gmc_ctx *c, int o, int n) {
  const int v = c->keep_derived ? 1 : 0;
  c->derived_kept[o] = v;
  if (c->cfg.time_scheme != GMC_TIME_EULER) {
    c->derived_kept[n] = v;
    c->derived_kept[3] = v;
  }
}

void derived_forget(gmc_ctx *c) {
  for (int &k : c->derived_kept) k = 0;
}

// gmc_step: replay a captured graph of `steps` fused steps from old level o
// (the levels swap after every step).  Kernels inside one graph follow each
// other with no gap, while consecutive graph launches leave one (about 9 us at
// 1440 x 721), so long runs replay graphs of several steps.
int fused_step_graph(gmc_ctx *c, int o, int n, double dt, int steps) {
  int rc = fused_prepare(c);
  if (rc) return rc;
  hipGraphExec_t exec = nullptr;
  for (auto &sg : c->graphs)
    if (sg.o == o && sg.n == n && sg.dt == dt && sg.keep == c->keep_derived && sg.steps == steps) exec = sg.exec;
  if (!exec) {
    hipGraph_t graph = nullptr;
    rc = check_hip(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    int rc2 = GMC_OK, a = o, b = n;
    for (int s = 0; s < steps && !rc2; s++) {
      rc2 = fused_integrate(c, a, b, dt);
      std::swap(a, b);
    }
    rc = check_hip(c, hipStreamEndCapture(c->stream, &graph), "end capture");
    if (rc2) return rc2;
    if (rc) return rc;
    rc = check_hip(c, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    (void)hipGraphDestroy(graph);
    if (rc) return rc;
    c->graphs.push_back({o, n, dt, c->keep_derived, steps, exec});
  }
  return check_hip(c, hipGraphLaunch(exec, c->stream), "graph launch");
}

}  // namespace gmc

// ===========================================================================
// Batched ensembles (include/gmcore_ens.h)
namespace gmc {
namespace {

// launch shapes the tuner chooses from: rows of the k_tend tile and the
// workgroups of the persistent tendency launch (0: one per tile, else
// mult x 256)
struct EnsShape {
  int ty, mult;
  const char *name;
};
constexpr EnsShape kEnsShapes[] = {
    {11, 0, "tile 58x11, one workgroup per tile"}, {7, 0, "tile 58x7, one workgroup per tile"},
    {5, 0, "tile 58x5, one workgroup per tile"},   {11, 4, "tile 58x11, 1024 persistent workgroups"},
    {7, 4, "tile 58x7, 1024 persistent workgroups"}, {5, 4, "tile 58x5, 1024 persistent workgroups"},
    {11, 8, "tile 58x11, 2048 persistent workgroups"}, {7, 8, "tile 58x7, 2048 persistent workgroups"},
    {5, 8, "tile 58x5, 2048 persistent workgroups"}};
constexpr int kNumEnsShapes = sizeof(kEnsShapes) / sizeof(kEnsShapes[0]);

int stage_count(int scheme) { return scheme == GMC_TIME_EULER ? 1 : 3; }

// (old, star, new) level of each stage and its dt, as fused_integrate
void stage_levels(int scheme, int o, int n, double dt, int st, int &so, int &ss, int &sn, double &sdt) {
  so = o;
  if (scheme == GMC_TIME_EULER) { ss = o; sn = n; sdt = dt; return; }
  if (st == 0) { ss = o; sn = n; sdt = scheme == GMC_TIME_WRFRK3 ? dt / 3.0 : dt / 2.0; }
  else if (st == 1) { ss = n; sn = 3; sdt = dt / 2.0; }
  else { ss = 3; sn = n; sdt = dt; }
}

}  // namespace
}  // namespace gmc

struct gmc_ens {
  std::vector<gmc_ctx *> m;
  int scheme = 0, nstage = 3;
  int shape = 0;
  gmc::EnsMember *d_mem = nullptr;  // [2 parities][nstage][members]
  gmc::Item *d_items = nullptr;
  int n_gz = 0, n_uv = 0;
  size_t lds_gz = 0, lds_uv = 0;
  int total_tiles = 0;
  std::vector<double> dt;           // dt the tables were built for
  int built_shape = -1;
  // graphs of one step and of kEnsBlock steps, by the parity they start from
  hipGraphExec_t graph[2] = {nullptr, nullptr}, graph_block[2] = {nullptr, nullptr};
  hipStream_t stream = nullptr;
  std::string err;
};

namespace gmc {
namespace {

int eerr(gmc_ens *e, int code, const std::string &msg) {
  if (e) e->err = msg;
  return code;
}
int ehip(gmc_ens *e, hipError_t x, const char *what) {
  if (x == hipSuccess) return GMC_OK;
  return eerr(e, x == hipErrorOutOfMemory ? GMC_ENOMEM : GMC_EDEVICE, std::string(what) + ": " + hipGetErrorString(x));
}

void drop_graphs(gmc_ens *e) {
  for (hipGraphExec_t *g : {&e->graph[0], &e->graph[1], &e->graph_block[0], &e->graph_block[1]})
    if (*g) { hip_note(hipGraphExecDestroy(*g), "ensemble graph destroy"); *g = nullptr; }
}

// descriptor tables for the current shape and dts
int build_tables(gmc_ens *e) {
  const int nm = (int)e->m.size(), ty = kEnsShapes[e->shape].ty;
  std::vector<EnsMember> tab((size_t)2 * e->nstage * nm);
  int tiles = 0;
  std::vector<int> tile0(nm), nx(nm);
  for (int k = 0; k < nm; k++) {
    const Grid &g = e->m[k]->grid;
    nx[k] = (g.nlon + TX - 1) / TX;
    tile0[k] = tiles;
    tiles += nx[k] * ((g.nlat + ty - 1) / ty);
  }
  e->total_tiles = tiles;
  for (int par = 0; par < 2; par++)
    for (int st = 0; st < e->nstage; st++)
      for (int k = 0; k < nm; k++) {
        gmc_ctx *c = e->m[k];
        const int o = par == 0 ? 1 : 2, n = par == 0 ? 2 : 1;
        int so, ss, sn;
        double sdt;
        stage_levels(e->scheme, o, n, e->dt[k], st, so, ss, sn, sdt);
        double **O = c->state[so], **S = c->state[ss], **N = c->state[sn];
        double *du = c->tend[so][T_DU], *dv = c->tend[so][T_DV], *gz_new = c->tend[so][T_DGZ];
        EnsMember &d = tab[((size_t)par * e->nstage + st) * nm + k];
        d.g = c->grid;
        d.M = Mesh{c->d_mesh, c->grid.nlat};
        d.mrows = c->d_mrows;
        d.big = c->big;
        d.small1 = c->small1;
        d.grav = c->cfg.g;
        d.rgrav = 1.0 / c->cfg.g;
        d.dt = sdt;
        d.wgt = c->cfg.upwind_wgt_pv;
        d.stokes = c->cfg.pv_pole_stokes;
        d.pv_scheme = c->cfg.pv_scheme;
        d.order = pv_order(c->cfg);
        d.coriolis = c->cfg.coriolis_scheme;
        d.sp = StagePtrs{S[S_GZ], c->gzs, S[S_U_F], S[S_V_F], O[S_GZ], gz_new, du, dv};
        d.up = UVPtrs{O[S_U_LON], O[S_V_LAT], du, dv, N[S_GZ_F], N[S_U_LON], N[S_V_LAT], N[S_U_F], N[S_V_F]};
        d.gz_f_out = N[S_GZ_F];
        d.gz_out = N[S_GZ];
        d.tile0 = tile0[k];
        d.nx = nx[k];
      }
  int rc = ehip(e, hipMemcpy(e->d_mem, tab.data(), sizeof(EnsMember) * tab.size(), hipMemcpyHostToDevice), "tables");
  if (rc) return rc;
  drop_graphs(e);
  e->built_shape = e->shape;
  return GMC_OK;
}

template <int TY>
void launch_tend(gmc_ens *e, const EnsMember *mem, int grid) {
  using S = Tile<TY>;
  hipLaunchKernelGGL(k_tend_ens<S>, dim3(grid), dim3(S::NT), 0, e->stream, mem, (int)e->m.size(), e->total_tiles);
}

// the launches of one ensemble step for parity par (0: old 1 -> new 2)
void ens_launch_step(gmc_ens *e, int par) {
  const int nm = (int)e->m.size();
  const EnsShape &sh = kEnsShapes[e->shape];
  // k_tend_ens covers every tile when grid == total or grid >= NXCD
  const int grid = sh.mult ? std::min(e->total_tiles, sh.mult * 256) : e->total_tiles;
  for (int st = 0; st < e->nstage; st++) {
    const EnsMember *mem = e->d_mem + ((size_t)par * e->nstage + st) * nm;
    switch (sh.ty) {
    case 5: launch_tend<5>(e, mem, grid); break;
    case 7: launch_tend<7>(e, mem, grid); break;
    default: launch_tend<11>(e, mem, grid);
    }
    hipLaunchKernelGGL(k_gzf_ens, dim3(e->n_gz), dim3(NF), e->lds_gz, e->stream, mem, e->d_items, e->n_gz);
    hipLaunchKernelGGL(k_uvf_ens, dim3(e->n_uv), dim3(NF), e->lds_uv, e->stream, mem, e->d_items + e->n_gz, e->n_uv);
  }
}

// steps per graph of a long run (even: the parity ends where it began); as
// gmc_step's, since consecutive graph launches leave a gap that kernels
// inside one graph do not
constexpr int kEnsBlock = 8;

// the graph of `steps` (1 or kEnsBlock) ensemble steps from parity par
int ens_graph(gmc_ens *e, int par, int steps, hipGraphExec_t *out) {
  hipGraphExec_t &slot = steps == 1 ? e->graph[par] : e->graph_block[par];
  if (slot) { *out = slot; return GMC_OK; }
  hipGraph_t graph = nullptr;
  int rc = ehip(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  for (int s = 0; s < steps; s++) ens_launch_step(e, (par + s) % 2);
  const hipError_t le = hipGetLastError();
  rc = ehip(e, hipStreamEndCapture(e->stream, &graph), "end capture");
  if (!rc) rc = ehip(e, le, "ensemble launch");
  if (!rc) rc = ehip(e, hipGraphInstantiate(&slot, graph, nullptr, nullptr, 0), "graph instantiate");
  if (graph) (void)hipGraphDestroy(graph);
  if (!rc) *out = slot;
  return rc;
}

int ens_run(gmc_ens *e, const double *dt, int nsteps) {
  const int nm = (int)e->m.size();
  for (gmc_ctx *c : e->m) {
    // the batch steps every member from member 0's time level
    if (c->old_level != e->m[0]->old_level)
      return eerr(e, GMC_ESTATE, "members are at different time levels (one was stepped on its own)");
  }
  for (gmc_ctx *c : e->m) derived_forget(c);
  bool same = e->built_shape == e->shape;
  for (int k = 0; k < nm && same; k++) same = e->dt[k] == dt[k];
  if (!same) {
    for (int k = 0; k < nm; k++) e->dt[k] = dt[k];
    int rc = build_tables(e);
    if (rc) return rc;
  }
  for (int s = 0; s < nsteps;) {
    const int par = e->m[0]->old_level == 1 ? 0 : 1;
    const int steps = nsteps - s >= kEnsBlock ? kEnsBlock : 1;
    hipGraphExec_t g = nullptr;
    int rc = ens_graph(e, par, steps, &g);
    if (!rc) rc = ehip(e, hipGraphLaunch(g, e->stream), "graph launch");
    if (rc) return rc;
    for (int k = 0; k < steps; k++)
      for (gmc_ctx *c : e->m) std::swap(c->old_level, c->new_level);
    s += steps;
  }
  return GMC_OK;
}

}  // namespace
}  // namespace gmc

using namespace gmc;

extern "C" {

int gmc_ens_create(gmc_ctx *const *members, int32_t n, gmc_ens **out) {
  if (!out) return GMC_EINVAL;
  *out = nullptr;
  if (!members || n < 1) return GMC_EINVAL;
  gmc_ens *e = new gmc_ens();
  auto fail = [&](int code, const char *msg) {
    gmc_ens_destroy(e);
    (void)msg;
    return code;
  };
  for (int k = 0; k < n; k++) {
    gmc_ctx *c = members[k];
    if (!c || !c->prepared || c->cfg.use_div_damp || c->cfg.time_scheme != members[0]->cfg.time_scheme ||
        c->old_level != members[0]->old_level || c->grid.nlon > Tile<5>::kMaxPoleLon)
      return fail(GMC_EINVAL, "member not batchable");
    for (int j = 0; j < k; j++)
      if (members[j] == c) return fail(GMC_EINVAL, "member listed twice");
    if (fused_prepare(c)) return fail(GMC_EINVAL, "member plan");
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(GMC_EDEVICE, "member sync");
    e->m.push_back(c);
  }
  e->scheme = members[0]->cfg.time_scheme;
  e->nstage = stage_count(e->scheme);
  e->dt.assign(n, 0.0);
  // filter work items of every member, tagged with the member index
  std::vector<Item> gz, uv;
  size_t lgz = 0, luv = 0;
  int maxb = 1, maxs = 1;
  for (int k = 0; k < n; k++) {
    gmc_ctx *c = e->m[k];
    const Grid &g = c->grid;
    std::vector<int32_t> nb_lon(g.nlat), nb_lat(g.nhalf), ns_lon(g.nlat);
    if (hipMemcpy(nb_lon.data(), c->big.ngrid_lon, sizeof(int32_t) * g.nlat, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(nb_lat.data(), c->big.ngrid_lat, sizeof(int32_t) * g.nhalf, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ns_lon.data(), c->small1.ngrid_lon, sizeof(int32_t) * g.nlat, hipMemcpyDeviceToHost) != hipSuccess)
      return fail(GMC_EDEVICE, "filter tables");
    std::vector<Item> a, b;
    for (int r = 0; r < g.nlat; r++)
      add_row(a, r, 0, nb_lon[r] + std::max(1, nb_lon[r] / 2), g.nlon, lgz, std::max(nb_lon[r], ns_lon[r]));
    for (int r = 0; r < g.nlat; r++) add_row(b, r, 1, nb_lon[r], g.nlon, luv, nb_lon[r]);
    for (int r = 0; r < g.nhalf; r++) add_row(b, r, 2, nb_lat[r], g.nlon, luv, nb_lat[r]);
    for (auto &it : a) {
      it.n = nb_lon[it.row];
      it.n2 = ns_lon[it.row];
      it.member = k;
      gz.push_back(it);
    }
    for (auto &it : b) {
      it.member = k;
      uv.push_back(it);
    }
    maxb = std::max(maxb, c->big.max_taps);
    maxs = std::max(maxs, c->small1.max_taps);
  }
  e->n_gz = (int)gz.size();
  e->n_uv = (int)uv.size();
  e->lds_gz = lgz + sizeof(double) * (maxb + maxs);
  e->lds_uv = luv + sizeof(double) * maxb;
  gz.insert(gz.end(), uv.begin(), uv.end());
  int rc = ehip(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "stream");
  if (!rc) rc = ehip(e, hipMalloc((void **)&e->d_items, sizeof(Item) * gz.size()), "items");
  if (!rc) rc = ehip(e, hipMemcpy(e->d_items, gz.data(), sizeof(Item) * gz.size(), hipMemcpyHostToDevice), "items");
  if (!rc) rc = ehip(e, hipMalloc((void **)&e->d_mem, sizeof(EnsMember) * 2 * e->nstage * n), "tables");
  if (!rc && e->lds_gz > 65536)
    rc = ehip(e, hipFuncSetAttribute((const void *)k_gzf_ens, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->lds_gz), "lds");
  if (!rc && e->lds_uv > 65536)
    rc = ehip(e, hipFuncSetAttribute((const void *)k_uvf_ens, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->lds_uv), "lds");
  if (!rc && std::max(e->lds_gz, e->lds_uv) > 160 * 1024) rc = eerr(e, GMC_EINVAL, "filter window exceeds LDS");
  if (rc) {
    gmc_ens_destroy(e);
    return rc;
  }
  *out = e;
  return GMC_OK;
}

int gmc_ens_destroy(gmc_ens *e) {
  if (!e) return GMC_EINVAL;
  if (e->stream) hip_note(hipStreamSynchronize(e->stream), "ensemble sync");
  drop_graphs(e);
  if (e->d_mem) hip_note(hipFree(e->d_mem), "ensemble tables free");
  if (e->d_items) hip_note(hipFree(e->d_items), "ensemble items free");
  if (e->stream) hip_note(hipStreamDestroy(e->stream), "ensemble stream destroy");
  delete e;
  return GMC_OK;
}

const char *gmc_ens_last_error(const gmc_ens *e) { return e ? e->err.c_str() : "null ensemble"; }

int gmc_ens_step(gmc_ens *e, const double *dt, int32_t nsteps) {
  if (!e) return GMC_EINVAL;
  if (!dt || nsteps < 0) return eerr(e, GMC_EINVAL, "bad step arguments");
  for (size_t k = 0; k < e->m.size(); k++)
    if (!(dt[k] > 0)) return eerr(e, GMC_EINVAL, "dt must be positive");
  int rc = ens_run(e, dt, nsteps);
  if (!rc) rc = ehip(e, hipStreamSynchronize(e->stream), "ensemble sync");
  return rc;
}

int gmc_ens_shape(gmc_ens *e, int32_t shape) {
  if (!e) return GMC_EINVAL;
  if (shape < 0 || shape >= kNumEnsShapes) return eerr(e, GMC_EINVAL, "no such shape");
  e->shape = shape;
  return GMC_OK;
}

const char *gmc_ens_shape_name(int32_t shape) {
  return shape >= 0 && shape < kNumEnsShapes ? kEnsShapes[shape].name : "";
}

int gmc_ens_autotune(gmc_ens *e, const double *dt, int32_t trial_steps, int32_t *choice, double *times_ms,
                     int32_t cap, int32_t *ncand) {
  if (!e) return GMC_EINVAL;
  if (!dt || trial_steps < 1 || !choice) return eerr(e, GMC_EINVAL, "bad autotune arguments");
  const int nm = (int)e->m.size();
  // snapshot every member's fields and levels into one buffer; the trials
  // run on the members and the snapshot is restored afterwards
  std::vector<size_t> off(nm + 1, 0);
  for (int k = 0; k < nm; k++) off[k + 1] = off[k] + (e->m[k]->slab_bytes + 255) / 256 * 256;
  char *save = nullptr;
  std::vector<int> levels(nm);
  for (int k = 0; k < nm; k++) levels[k] = e->m[k]->old_level;
  int rc = GMC_OK;
  for (int k = 0; k < nm && !rc; k++) rc = ehip(e, hipStreamSynchronize(e->m[k]->stream), "member sync");
  if (!rc) rc = ehip(e, hipMalloc((void **)&save, off[nm]), "snapshot alloc");
  for (int k = 0; k < nm && !rc; k++)
    rc = ehip(e, hipMemcpyAsync(save + off[k], e->m[k]->slab, e->m[k]->slab_bytes, hipMemcpyDeviceToDevice, e->stream),
              "snapshot");
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (!rc) rc = ehip(e, hipEventCreate(&ev0), "event");
  if (!rc) rc = ehip(e, hipEventCreate(&ev1), "event");
  const int keep = e->shape;
  int best = keep;
  float best_ms = 0.f;
  for (int s = 0; s < kNumEnsShapes && !rc; s++) {
    e->shape = s;
    rc = ens_run(e, dt, 1);  // warm-up, builds tables and graphs
    if (!rc) rc = ehip(e, hipEventRecord(ev0, e->stream), "event");
    if (!rc) rc = ens_run(e, dt, trial_steps);
    if (!rc) rc = ehip(e, hipEventRecord(ev1, e->stream), "event");
    if (!rc) rc = ehip(e, hipEventSynchronize(ev1), "event sync");
    float ms = 0.f;
    if (!rc) rc = ehip(e, hipEventElapsedTime(&ms, ev0, ev1), "elapsed");
    ms /= trial_steps;
    if (!rc && times_ms && s < cap) times_ms[s] = ms;
    if (!rc && (s == 0 || ms < best_ms)) { best_ms = ms; best = s; }
  }
  if (save) {
    for (int k = 0; k < nm; k++) {
      if (!rc)
        rc = ehip(e, hipMemcpyAsync(e->m[k]->slab, save + off[k], e->m[k]->slab_bytes, hipMemcpyDeviceToDevice,
                                    e->stream), "restore");
      e->m[k]->old_level = levels[k];
      e->m[k]->new_level = levels[k] == 1 ? 2 : 1;
    }
    if (!rc) rc = ehip(e, hipStreamSynchronize(e->stream), "restore sync");
    else hip_note(hipStreamSynchronize(e->stream), "autotune sync");
    hip_note(hipFree(save), "snapshot free");
  }
  if (ev0) hip_note(hipEventDestroy(ev0), "event destroy");
  if (ev1) hip_note(hipEventDestroy(ev1), "event destroy");
  e->shape = rc ? keep : best;
  if (!rc) {
    *choice = best;
    if (ncand) *ncand = kNumEnsShapes;
  }
  return rc;
}

int gmc_ens_stream(gmc_ens *e, void **stream) {
  if (!e || !stream) return GMC_EINVAL;
  *stream = (void *)e->stream;
  return GMC_OK;
}

}  // extern "C"
