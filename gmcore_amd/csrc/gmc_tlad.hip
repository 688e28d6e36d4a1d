// gmc_tlad.hip -- tangent-linear (TL) and adjoint (AD) of the shallow-water
// step on gfx950: the C ABI of include/gmcore_tlad.h.
//
// The step is time_integrator (pc2 / wrfrk3 / euler) of stages
// step_forward_backward (time_schemes_mod.F90:102-118, 244-288); a stage maps
// (old state o, star state s) to the new state n.  The state seen by a stage
// is (u, v, gz) of o and (gz, u_f, v_f) of s, with u_f = BF(u), v_f = BF(v)
// formed by the previous stage (BF the big zonal filter).  The stage's
// trajectory ("tape") is the set of intermediates of operators_prepare in
// HBM: the TL step takes it from the operator-level kernels
// (ops_stage_tape), the adjoint from the fused tendency kernel's taping
// variant (fused_stage_tape, same bits); the TL kernels differentiate each
// operator around it, and the AD kernels apply the transposes in reverse
// order.
//
// Every adjoint kernel is a gather: one thread owns one adjoint output and
// adds the contributions of every forward output that read it, so no atomics
// are needed and the result is deterministic.  The zonal filters have
// symmetric weights, so their transpose is the filter itself.  Non-smooth
// points (|t| and sign(t) of the upwind PV weights) take the derivative of
// the branch the trajectory took.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "gmc_internal.hpp"
#include "gmc_point.hpp"
#include "../../include/gmcore_tlad.h"

namespace gmc {
namespace {

constexpr int BX = 64, BY = 4;
dim3 grid2d(const Grid &g, int rows) { return dim3((g.nlon + BX - 1) / BX, (rows + BY - 1) / BY); }

// perturbation / adjoint fields of one state level
enum { X_U, X_V, X_GZ, X_UF, X_VF, X_N };
// stage work fields (perturbations in TL, adjoints in AD)
enum {
  W_M, W_MLON, W_MLAT, W_MVTX, W_MFX, W_MFY, W_MFXLAT, W_MFYLON, W_ULAT, W_VLON, W_KE, W_VOR, W_PV, W_PVLON,
  W_PVLAT, W_QHU, W_QHV, W_DU, W_DV, W_DGZ, W_GZNEW, W_GZF, W_T1, W_T2, W_N
};

// the forward trajectory of one stage (fields of the star state, its tend);
// the adjoint's tapes also hold the upwind-pv derivatives (UpwD dt, dn,
// df[0..3]) of the pv_lon edges (dlon) and the pv_lat edges (dlat)
struct Tape {
  const double *m_lon, *m_lat, *m_vtx, *mfx, *mfy, *mfx_lat, *mfy_lon, *u_lat, *v_lon, *pv, *pv_lon, *pv_lat;
  const double *u_f, *v_f;
  const double *dlon[6], *dlat[6];
};

struct Opt {
  double grav;
  int pv_scheme, order, coriolis, stokes;
  double wgt;
};

// derivatives of pv_upwind_edge on the branch the trajectory took
__device__ __forceinline__ UpwD upwind_d(const Opt &o, double t, double n, int lat_edge, double f1, double f2,
                                         double f3, double f4) {
  return upwind_deriv(o.pv_scheme, o.order, o.wgt, t, n, lat_edge, f1, f2, f3, f4);
}

// ===========================================================================
// Tangent-linear kernels.  x_o: perturbation of the old state, x_s: of the
// star state, W: stage work perturbations, x_n: of the new state.
struct TLPtrs {
  const double *o_u, *o_v, *o_gz, *s_gz, *s_uf, *s_vf;
  double *w[W_N];
  double *n_u, *n_v, *n_gz;
  double *pole;  // [0..1] ke, [2..3] dmfdlat, [4..5] vor: perturbations at the poles
};

__global__ void t_m(Grid g, Mesh M, Opt o, TLPtrs P) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch;
  const int ip = wrap(i + 1, g.nlon);
  auto dm = [&](int rr, int ii) { return P.s_gz[rr * p + ii] / o.grav; };
  P.w[W_M][r * p + i] = dm(r, i);
  if (r >= 1 && r <= g.nlat - 2) P.w[W_MLON][r * p + i] = (dm(r, i) + dm(r, ip)) * 0.5;
  if (r <= g.nhalf - 1) {
    P.w[W_MLAT][r * p + i] = (dm(r, i) + dm(r + 1, i)) * 0.5;
    P.w[W_MVTX][r * p + i] = m_vtx_of(dm(r, i), dm(r, ip), dm(r + 1, i), dm(r + 1, ip), M, r);
  }
}

__global__ void t_mf(Grid g, Mesh M, Tape T, TLPtrs P) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch;
  const int im = wrap(i - 1, g.nlon), ip = wrap(i + 1, g.nlon);
  auto dmfx = [&](int rr, int ii) {
    const int64_t k = rr * p + ii;
    return (rr == 0 || rr == g.nlat - 1) ? 0.0 : P.w[W_MLON][k] * T.u_f[k] + T.m_lon[k] * P.s_uf[k];
  };
  auto dmfy = [&](int rr, int ii) {
    const int64_t k = rr * p + ii;
    return P.w[W_MLAT][k] * T.v_f[k] + T.m_lat[k] * P.s_vf[k];
  };
  const int64_t k = r * p + i;
  P.w[W_MFX][k] = dmfx(r, i);
  if (r >= 1 && r <= g.nlat - 2) {
    const double d = tangent_sum(M(GMC_FULL_TANGENT_WGT1, r), M(GMC_FULL_TANGENT_WGT2, r), dmfy(r - 1, i),
                                 dmfy(r - 1, ip), dmfy(r, i), dmfy(r, ip));
    P.w[W_MFYLON][k] = d;
    P.w[W_VLON][k] = (d - T.v_lon[k] * P.w[W_MLON][k]) / T.m_lon[k];
  }
  if (r <= g.nhalf - 1) {
    P.w[W_MFY][k] = dmfy(r, i);
    const double d = tangent_sum(M(GMC_HALF_TANGENT_WGT1, r), M(GMC_HALF_TANGENT_WGT2, r), dmfx(r, im),
                                 dmfx(r, i), dmfx(r + 1, im), dmfx(r + 1, i));
    P.w[W_MFXLAT][k] = d;
    P.w[W_ULAT][k] = (d - T.u_lat[k] * P.w[W_MLAT][k]) / T.m_lat[k];
  }
}

// pole perturbations (one wavefront per pole), as k_pole_sums
__global__ void t_pole(Grid g, Mesh M, Tape T, TLPtrs P) {
  const int north = blockIdx.x, lane = threadIdx.x;
  const int vr = north ? g.nhalf - 1 : 0, cr = north ? g.nlat - 1 : 0;
  const int64_t p = g.pitch;
  const double *v = T.v_f + vr * p, *dv = P.s_vf + vr * p, *dmfy = P.w[W_MFY] + vr * p, *dul = P.w[W_ULAT] + vr * p;
  const double le = M(GMC_LE_LAT, vr), ac = M(GMC_AREA_CELL, cr), nl = (double)g.nlon;
  const double ske = wave_zonal_sum(g.nlon, lane, [&](int i) { return 2 * v[i] * dv[i]; });
  const double smf = wave_zonal_sum(g.nlon, lane, [&](int i) { return north ? -dmfy[i] : dmfy[i]; });
  const double svo = wave_zonal_sum(g.nlon, lane, [&](int i) { return north ? dul[i] * le : -dul[i] * le; });
  if (lane == 0) {
    P.pole[north] = ske / nl;
    P.pole[2 + north] = smf * le / nl / ac;
    P.pole[4 + north] = svo / nl / ac;
  }
}

__global__ void t_ke_pv(Grid g, Mesh M, Opt o, Tape T, TLPtrs P) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  const int im = wrap(i - 1, g.nlon), ip = wrap(i + 1, g.nlon);
  // ke
  if (r == 0 || r == g.nlat - 1) {
    P.w[W_KE][k] = P.pole[r == 0 ? 0 : 1];
  } else {
    const double uw = T.u_f[r * p + im], ue = T.u_f[k], vs = T.v_f[(r - 1) * p + i], vn = T.v_f[k];
    P.w[W_KE][k] = (M(GMC_AREA_LON_WEST, r) * (2 * uw * P.s_uf[r * p + im]) +
                    M(GMC_AREA_LON_EAST, r) * (2 * ue * P.s_uf[k]) +
                    M(GMC_AREA_LAT_NORTH, r - 1) * (2 * vs * P.s_vf[(r - 1) * p + i]) +
                    M(GMC_AREA_LAT_SOUTH, r) * (2 * vn * P.s_vf[k])) /
                   M(GMC_AREA_CELL, r);
  }
  // vor, pv
  if (r <= g.nhalf - 1) {
    double dw = vor_of(P.s_uf[k], P.s_uf[k + p], P.s_vf[r * p + ip], P.s_vf[k], M, r);
    if (o.stokes && r == 0) dw = P.pole[4];
    else if (o.stokes && r == g.nhalf - 1) dw = P.pole[5];
    P.w[W_VOR][k] = dw;
    P.w[W_PV][k] = (dw - T.pv[k] * P.w[W_MVTX][k]) / T.m_vtx[k];
  }
}

__global__ void t_interp(Grid g, Opt o, Tape T, TLPtrs P) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  auto PV = [&](const double *a, int rr, int ii) { mirror_half(rr, ii, g); return a[rr * p + ii]; };
  const double *dpv = P.w[W_PV];
  if (r == 0 || r == g.nlat - 1) P.w[W_PVLON][k] = 0.0;   // pv_lon is not formed on the pole rows
  if (r >= 1 && r <= g.nlat - 2) {
    const UpwD d = upwind_d(o, T.v_lon[k], T.u_f[k], 0, PV(T.pv, r - 2, i), PV(T.pv, r - 1, i), PV(T.pv, r, i),
                            PV(T.pv, r + 1, i));
    P.w[W_PVLON][k] = d.dt * P.w[W_VLON][k] + d.dn * P.s_uf[k] + d.df[0] * PV(dpv, r - 2, i) +
                      d.df[1] * PV(dpv, r - 1, i) + d.df[2] * PV(dpv, r, i) + d.df[3] * PV(dpv, r + 1, i);
  }
  if (r <= g.nhalf - 1) {
    const int i1 = wrap(i - 2, g.nlon), i2 = wrap(i - 1, g.nlon), i4 = wrap(i + 1, g.nlon);
    const UpwD d = upwind_d(o, T.u_lat[k], T.v_f[k], 1, T.pv[r * p + i1], T.pv[r * p + i2], T.pv[k], T.pv[r * p + i4]);
    P.w[W_PVLAT][k] = d.dt * P.w[W_ULAT][k] + d.dn * P.s_vf[k] + d.df[0] * dpv[r * p + i1] +
                      d.df[1] * dpv[r * p + i2] + d.df[2] * dpv[k] + d.df[3] * dpv[r * p + i4];
  }
}

// grad_mf, Coriolis, grad_ke and the forward tendencies (perturbations)
__global__ void t_tend(Grid g, Mesh M, Opt o, Tape T, TLPtrs P) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  const int im = wrap(i - 1, g.nlon), ip = wrap(i + 1, g.nlon);
  const double *dmfx = P.w[W_MFX], *dmfy = P.w[W_MFY], *dpl = P.w[W_PVLON], *dpt = P.w[W_PVLAT], *dke = P.w[W_KE];
  // dmfdlon, dmfdlat
  double dlon = 0.0, dlat;
  if (r == 0) dlat = P.pole[2];
  else if (r == g.nlat - 1) dlat = P.pole[3];
  else {
    dlon = dmfdlon_of(dmfx[k], dmfx[r * p + im], M, r);
    dlat = dmfdlat_of(dmfy[k], dmfy[(r - 1) * p + i], M, r);
  }
  P.w[W_DGZ][k] = -(dlon + dlat) * o.grav;
  if (r <= g.nhalf - 1) {
    double dq;
    if (o.coriolis == 1) {
      const double w1 = M(GMC_HALF_TANGENT_WGT1, r), w2 = M(GMC_HALF_TANGENT_WGT2, r);
      const int64_t a = r * p + im, b = k, c = (r + 1) * p + im, d = (r + 1) * p + i;
      const double pt = T.pv_lat[k], dp = dpt[k];
      dq = (w1 * (dmfx[a] * (pt + T.pv_lon[a]) + T.mfx[a] * (dp + dpl[a]) + dmfx[b] * (pt + T.pv_lon[b]) +
                  T.mfx[b] * (dp + dpl[b])) +
            w2 * (dmfx[c] * (pt + T.pv_lon[c]) + T.mfx[c] * (dp + dpl[c]) + dmfx[d] * (pt + T.pv_lon[d]) +
                  T.mfx[d] * (dp + dpl[d]))) *
           0.5;
    } else {
      dq = P.w[W_MFXLAT][k] * T.pv_lat[k] + T.mfx_lat[k] * dpt[k];
    }
    P.w[W_QHU][k] = dq;
    P.w[W_DV][k] = -dq - (dke[k + p] - dke[k]) / M(GMC_DE_LAT, r);
  }
  if (r >= 1 && r <= g.nlat - 2) {
    double dq;
    if (o.coriolis == 1) {
      const double w1 = M(GMC_FULL_TANGENT_WGT1, r), w2 = M(GMC_FULL_TANGENT_WGT2, r);
      const int64_t a = (r - 1) * p + i, b = (r - 1) * p + ip, c = k, d = r * p + ip;
      const double pl = T.pv_lon[k], dp = dpl[k];
      dq = (w1 * (dmfy[a] * (pl + T.pv_lat[a]) + T.mfy[a] * (dp + dpt[a]) + dmfy[b] * (pl + T.pv_lat[b]) +
                  T.mfy[b] * (dp + dpt[b])) +
            w2 * (dmfy[c] * (pl + T.pv_lat[c]) + T.mfy[c] * (dp + dpt[c]) + dmfy[d] * (pl + T.pv_lat[d]) +
                  T.mfy[d] * (dp + dpt[d]))) *
           0.5;
    } else {
      dq = P.w[W_MFYLON][k] * T.pv_lon[k] + T.mfy_lon[k] * dpl[k];
    }
    P.w[W_QHV][k] = dq;
    P.w[W_DU][k] = dq - (dke[r * p + ip] - dke[k]) / M(GMC_DE_LON, r);
  }
}

__global__ void t_gznew(Grid g, double h, TLPtrs P) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t k = r * g.pitch + i;
  P.w[W_GZNEW][k] = P.o_gz[k] + h * P.w[W_DGZ][k];
}

// u, v update with the backward pass (pgf of the new gz_f) and the pole-row
// v smoothing
__global__ void t_uv(Grid g, Mesh M, double h, TLPtrs P) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  const double *gf = P.w[W_GZF];
  if (r >= 1 && r <= g.nlat - 2)
    P.n_u[k] = P.o_u[k] + h * (P.w[W_DU][k] - pgf_lon_of(gf[r * p + wrap(i + 1, g.nlon)], gf[k], M, r));
  if (r <= g.nhalf - 1) {
    auto vs = [&](int rr) {
      const int64_t q = rr * p + i;
      return P.o_v[q] + h * (P.w[W_DV][q] - pgf_lat_of(gf[q + p], gf[q], M, rr));
    };
    double v = vs(r);
    if (r == 0) v = pole_smooth(v, vs(1));
    else if (r == g.nhalf - 1) v = pole_smooth(v, vs(r - 1));
    P.n_v[k] = v;
  }
}

// ===========================================================================
// Adjoint kernels (gathers).  lam_n*: adjoint of the new state; the kernels
// add into the old state's (o_*) and the star state's (s_*) adjoints.
struct ADPtrs {
  const double *n_u, *n_v;
  double *n_gz, *n_uf, *n_vf;   // cleared by a_cori once the row kernels have read them
  double *o_u, *o_v, *o_gz, *s_gz, *s_uf, *s_vf;
  double *w[W_N];
  double *pole;  // [0..1] ke, [2..3] dmfdlat, [4..5] vor: sums of pole-row adjoints
};

// Work items of the adjoint's row kernels: a u row, a group of v rows
// (rows r0..r1, see v_group) or a gz row, over columns [c0, c1).
enum { RI_U, RI_V, RI_GZ };
struct RowItem {
  int32_t kind, r0, r1, c0, c1, pad;
};

// The metric table and its row records, which also hold the reciprocals of
// the divisor fields (area_vtx, area_cell, de_lon, de_lat): the adjoint
// kernels multiply by rcp(f, r) where a transpose divides by a metric term.
struct MeshR {
  const double *m, *rows;
  int nlat;
  __device__ __forceinline__ double operator()(int f, int r) const { return m[f * nlat + r]; }
  __device__ __forceinline__ double rcp(int f, int r) const {
    return rows[(int64_t)(r + 1) * kRowRec + GMC_MESH_NFIELDS + recip_slot(f)];
  }
};

// The transpose of m_lon, m_lat, m_vtx and m = (gz - gzs) / g at cell (R,
// i): the gz adjoint that lam m_lon (ll), lam m_lat (lt), lam m_vtx (lv)
// give the star state (a_m adds it; a_gzrow of the next stage adds it to
// the gz adjoint it stages, when the stage before it left it pending).
__device__ __forceinline__ double am_point(const Grid &g, const MeshR &M, double grav, const double *ll,
                                           const double *lt, const double *lv, int R, int i) {
  const int64_t p = g.pitch, k = R * p + i;
  const int im = wrap(i - 1, g.nlon);
  double s = 0.0;
  if (R >= 1 && R <= g.nlat - 2) s += 0.5 * (ll[k] + ll[R * p + im]);
  if (R <= g.nhalf - 1) {
    s += 0.5 * lt[k];
    s += (lv[k] + lv[R * p + im]) * M(GMC_AREA_SUBCELL2, R) * M.rcp(GMC_AREA_VTX, R);
  }
  if (R >= 1) {
    s += 0.5 * lt[k - p];
    s += (lv[k - p] + lv[(R - 1) * p + im]) * M(GMC_AREA_SUBCELL1, R) * M.rcp(GMC_AREA_VTX, R - 1);
  }
  return s / grav;
}

// Zonal-filter staging: slot[t] = x[(c0 - hn + t) mod nlon] for t in [0, w +
// 2 hn), so the taps of column c0 + j start at slot + j; a row's n weights
// are staged next to it.
constexpr int NR = 512;   // threads of the row kernels
__device__ __forceinline__ void stage_window(double *slot, const double *x, int nlon, int c0, int w, int hn) {
  for (int t = threadIdx.x; t < w + 2 * hn; t += NR) {
    int c = (c0 - hn + t) % nlon;
    if (c < 0) c += nlon;
    slot[t] = x[c];
  }
}
__device__ __forceinline__ void stage_weights(double *dst, const double *src, int n) {
  for (int k = threadIdx.x; k < n; k += NR) dst[k] = src[k];
}

// v rows that share a work item of a_uvrow: the pole smoothing's transpose
// couples v rows 0 and 1, and nhalf-2 and nhalf-1
constexpr int kMaxVGroup = 3;
__host__ __device__ inline int v_groups(int nhalf) { return nhalf < 4 ? 1 : nhalf - 2; }
__host__ __device__ inline void v_group(int q, int nhalf, int &ra, int &rb) {
  if (nhalf < 4) { ra = 0; rb = nhalf - 1; }
  else if (q == 0) { ra = 0; rb = 1; }
  else if (q == nhalf - 3) { ra = nhalf - 2; rb = nhalf - 1; }
  else { ra = rb = q + 1; }
}

// Adjoint of the u, v update with its filters, one workgroup per work item
// (a u row, or a group of v rows, over a column chunk).  With lam u_f, lam
// v_f the adjoints of u_f = BF(u), v_f = BF(v) of the new state (BF
// symmetric, so its transpose is itself): lam u* = BF(lam u_f) + lam u, lam
// v* likewise, then the transpose of the pole-row smoothing (v(0) = 0.4 v(0)
// + 0.6 v(1), v(nhalf-1) likewise); lam_o u, v += lam u*, lam v* and W_DU,
// W_DV = h lam u*, h lam v* (the adjoints of the tendencies, which enter the
// update as h du, h dv).  The new state's u, v adjoints are cleared once
// read; its u_f, v_f adjoints, which neighbouring chunks read as halo, are
// cleared later in the stage (a_cori).  LDS: a u item takes its weights and
// window; a v group its rows' weights, windows and lam v* rows.
struct UVRow {
  double *n_u, *n_v;
  const double *n_uf, *n_vf;
  double *o_u, *o_v, *du, *dv;
};
__global__ void __launch_bounds__(NR) a_uvrow(Grid g, FilterDev f, double h, const RowItem *__restrict__ items,
                                              UVRow A) {
  extern __shared__ double xs[];
  const RowItem it = items[blockIdx.x];
  const int64_t p = g.pitch;
  const int c0 = it.c0, w = it.c1 - it.c0;
  if (it.kind == RI_U) {
    const int r = it.r0;
    double *__restrict__ u = A.n_u + r * p;
    double *__restrict__ ou = A.o_u + r * p;
    double *__restrict__ du = A.du + r * p;
    if (r == 0 || r == g.nlat - 1) {   // u is not formed on the pole rows
      for (int j = threadIdx.x; j < w; j += NR) u[c0 + j] = 0.0;
      return;
    }
    const int n = f.ngrid_lon[r], hn = n > 1 ? (n - 1) / 2 : 0;
    double *wt = xs, *x = xs + f.max_taps;
    // the first output's addends are loaded ahead of the barrier
    const int j0 = threadIdx.x;
    const double un0 = j0 < w ? u[c0 + j0] : 0.0, uo0 = j0 < w ? ou[c0 + j0] : 0.0;
    stage_weights(wt, f.wgt_lon + (int64_t)r * f.max_taps, n);
    stage_window(x, A.n_uf + r * p, g.nlon, c0, w, hn);
    __syncthreads();
    for (int j = threadIdx.x; j < w; j += NR) {
      const int c = c0 + j;
      const double un = j == j0 ? un0 : u[c], uo = j == j0 ? uo0 : ou[c];
      const double lu = (n > 1 ? filter_point(wt, n, x + j) : x[j]) + un;
      ou[c] = uo + lu;
      du[c] = h * lu;
      u[c] = 0.0;
    }
    return;
  }
  const int ra = it.r0, rb = it.r1, G = rb - ra + 1, span = w + f.max_taps;
  double *wv = xs, *xv = wv + G * f.max_taps, *lvs = xv + G * span;
  const int q0 = threadIdx.x;
  const int64_t k0 = q0 < G * w ? (ra + q0 / w) * p + c0 + q0 % w : 0;
  const double vn0 = q0 < G * w ? A.n_v[k0] : 0.0, vo0 = q0 < G * w ? A.o_v[k0] : 0.0;
  for (int j = 0; j < G; j++) {
    const int n = f.ngrid_lat[ra + j];
    stage_weights(wv + j * f.max_taps, f.wgt_lat + (int64_t)(ra + j) * f.max_taps, n);
    stage_window(xv + j * span, A.n_vf + (ra + j) * p, g.nlon, c0, w, n > 1 ? (n - 1) / 2 : 0);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < G * w; q += NR) {
    const int j = q / w, jc = q - j * w, r = ra + j, n = f.ngrid_lat[r];
    const double vn = q == q0 ? vn0 : A.n_v[r * p + c0 + jc];
    const double *x = xv + j * span + jc;
    lvs[q] = (n > 1 ? filter_point(wv + j * f.max_taps, n, x) : x[0]) + vn;
  }
  __syncthreads();
  const int last = g.nhalf - 1;
  const double wgt = 0.6;
  for (int q = threadIdx.x; q < G * w; q += NR) {
    const int j = q / w, r = ra + j;
    const int64_t k = r * p + c0 + (q - j * w);
    double ls = (r == 0 || r == last) ? (1 - wgt) * lvs[q] : lvs[q];
    if (r == 1) ls += wgt * lvs[q - w];          // row 0 is in the group
    if (r == last - 1) ls += wgt * lvs[q + w];   // and row nhalf-1
    A.o_v[k] = (q == q0 ? vo0 : A.o_v[k]) + ls;
    A.dv[k] = h * ls;
    A.n_v[k] = 0.0;
  }
}

// Adjoint of gz_f = BF(gz_new), gz = SF(gz_new) of the new state, of the pgf
// and of grad_ke, one workgroup per work item (a row over a column chunk;
// the pole rows whole).  lam gz_f comes from the pgf terms (pgf_lon =
// (gz_f(i+1) - gz_f(i)) / de_lon with lam pgf_lon = -lam du, pgf_lat
// likewise); lam ke has the same form (lam dkedlon = -lam du) and is kept in
// W_KE.  lam gz_new = BF(lam gz_f) + SF(lam gz); lam_o gz += lam gz_new,
// W_DGZ = -(h lam gz_new) g (= lam dmfdlon = lam dmfdlat).  The pole rows
// leave the sums of W_KE and W_DGZ over the row in pole[0..3].  The new
// state's gz adjoint is cleared later in the stage (a_cori).  When the stage
// before left its a_m pending, the gz adjoint staged here is that of the
// new state plus am_point of the previous stage's lam m_lon, m_lat, m_vtx
// (the same sum a_m would have stored).
struct GZRow {
  const double *du, *dv, *n_gz;
  double *o_gz, *ke, *dgz, *pole;
  const double *mlon, *mlat, *mvtx;   // with pending
  int pending;
};
__global__ void __launch_bounds__(NR) a_gzrow(Grid g, MeshR M, FilterDev fb, FilterDev fs, double h, double grav,
                                              const RowItem *__restrict__ items, GZRow A) {
  extern __shared__ double xs[];
  const RowItem it = items[blockIdx.x];
  const int r = it.r0, c0 = it.c0, w = it.c1 - it.c0;
  const int64_t p = g.pitch, o = r * p;
  // LDS: big weights | small weights | lam gz_f window | lam gz window | W_DGZ of a pole row
  double *wb = xs, *ws = wb + fb.max_taps, *xb = ws + fs.max_taps, *xg = xb + w + fb.max_taps;
  double *xd = xg + w + fs.max_taps;
  const int nb = fb.ngrid_lon[r], hb = nb > 1 ? (nb - 1) / 2 : 0;
  const int ns = fs.ngrid_lon[r], hs = ns > 1 ? (ns - 1) / 2 : 0;
  const bool u_row = r >= 1 && r <= g.nlat - 2;
  const int j0 = threadIdx.x;
  const double og0 = j0 < w ? A.o_gz[o + c0 + j0] : 0.0;   // ahead of the barrier
  stage_weights(wb, fb.wgt_lon + (int64_t)r * fb.max_taps, nb);
  stage_weights(ws, fs.wgt_lon + (int64_t)r * fs.max_taps, ns);
  for (int t = threadIdx.x; t < w + 2 * hb; t += NR) {
    int c = (c0 - hb + t) % g.nlon;
    if (c < 0) c += g.nlon;
    const int cm = c == 0 ? g.nlon - 1 : c - 1;
    double s = 0.0;
    if (u_row) s += (A.du[o + c] - A.du[o + cm]) * M.rcp(GMC_DE_LON, r);
    if (r >= 1) s += -A.dv[o - p + c] * M.rcp(GMC_DE_LAT, r - 1);
    if (r <= g.nhalf - 1) s += A.dv[o + c] * M.rcp(GMC_DE_LAT, r);
    xb[t] = s;
    if (t >= hb && t < hb + w) A.ke[o + c] = s;
  }
  if (A.pending) {
    for (int t = threadIdx.x; t < w + 2 * hs; t += NR) {
      int c = (c0 - hs + t) % g.nlon;
      if (c < 0) c += g.nlon;
      xg[t] = A.n_gz[o + c] + am_point(g, M, grav, A.mlon, A.mlat, A.mvtx, r, c);
    }
  } else {
    stage_window(xg, A.n_gz + o, g.nlon, c0, w, hs);
  }
  __syncthreads();
  const bool pole = r == 0 || r == g.nlat - 1;   // whole rows
  for (int j = threadIdx.x; j < w; j += NR) {
    const int c = c0 + j;
    const double og = j == j0 ? og0 : A.o_gz[o + c];
    const double lg =
        (nb > 1 ? filter_point(wb, nb, xb + j) : xb[j]) + (ns > 1 ? filter_point(ws, ns, xg + j) : xg[j]);
    A.o_gz[o + c] = og + lg;
    const double d = -(h * lg) * grav;
    A.dgz[o + c] = d;
    if (pole) xd[j] = d;
  }
  if (!pole) return;   // the same for the whole workgroup
  __syncthreads();
  if (threadIdx.x < kWave) {
    const int lane = threadIdx.x, north = r != 0;
    const double sk = wave_zonal_sum(g.nlon, lane, [&](int i) { return xb[hb + i]; });
    const double sd = wave_zonal_sum(g.nlon, lane, [&](int i) { return xd[i]; });
    if (lane == 0) {
      A.pole[north] = sk;
      A.pole[2 + north] = sd;
    }
  }
}

// Coriolis and grad_mf transposed: lam mfx, lam pv_lon (U rows), lam mfy,
// lam pv_lat (V rows); lam qhv = W_DU / h ... carried as lam du (= h lam u*):
// the forward tendencies enter the update as h du, so W_DU already holds
// lam du.
__global__ void a_cori(Grid g, MeshR M, Opt o, Tape T, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  const bool u_part = blockIdx.z == 0;   // the U points' and the V points' outputs in workgroups of their own
  // the new state's u_f, v_f, gz adjoints, read by the row kernels, are consumed
  if (u_part) {
    A.n_uf[k] = 0.0;
    A.n_gz[k] = 0.0;
  } else if (r < g.nhalf) {
    A.n_vf[k] = 0.0;
  }
  const int im = wrap(i - 1, g.nlon), ip = wrap(i + 1, g.nlon);
  const double *lqv = A.w[W_DU];   // lam qhv = lam du
  const double *ldv = A.w[W_DV];   // lam qhu = -lam dv
  const double *ldm = A.w[W_DGZ];  // lam dmfdlon = lam dmfdlat
  auto lqhu = [&](int rr, int ii) { return -ldv[rr * p + ii]; };
  if (u_part && r >= 1 && r <= g.nlat - 2) {
    // lam mfx: grad_mf
    double s = (ldm[k] - ldm[r * p + ip]) * M(GMC_LE_LON, r) * M.rcp(GMC_AREA_CELL, r);
    double spl = 0.0;
    if (o.coriolis == 1) {
      const double pl = T.pv_lon[k], mf = T.mfx[k];
      if (r <= g.nhalf - 1) {
        const double w1 = M(GMC_HALF_TANGENT_WGT1, r) * 0.5;
        const double qa = lqhu(r, ip), qb = lqhu(r, i);
        s += w1 * (qa * (T.pv_lat[r * p + ip] + pl) + qb * (T.pv_lat[k] + pl));
        spl += w1 * (qa + qb) * mf;
      }
      if (r - 1 >= 0) {
        const double w2 = M(GMC_HALF_TANGENT_WGT2, r - 1) * 0.5;
        const double qa = lqhu(r - 1, ip), qb = lqhu(r - 1, i);
        s += w2 * (qa * (T.pv_lat[(r - 1) * p + ip] + pl) + qb * (T.pv_lat[(r - 1) * p + i] + pl));
        spl += w2 * (qa + qb) * mf;
      }
      spl += lqv[k] * 0.5 *
             (M(GMC_FULL_TANGENT_WGT1, r) * (T.mfy[(r - 1) * p + i] + T.mfy[(r - 1) * p + ip]) +
              M(GMC_FULL_TANGENT_WGT2, r) * (T.mfy[k] + T.mfy[r * p + ip]));
    } else {
      spl += lqv[k] * T.mfy_lon[k];
    }
    A.w[W_MFX][k] = s;
    A.w[W_PVLON][k] = spl;
    // v_lon = mfy_lon / m_lon transposed, with the upwind weight's dependence on v_lon
    const double lvl = spl * T.dlon[0][k];
    double lmf = lvl / T.m_lon[k];
    if (o.coriolis != 1) lmf += lqv[k] * T.pv_lon[k];
    A.w[W_VLON][k] = lvl;
    A.w[W_MFYLON][k] = lmf;
  }
  if (!u_part && r <= g.nhalf - 1) {
    // lam mfy: grad_mf (cell rows r and r + 1) and the pole sums
    double s = 0.0;
    if (r >= 1) s += ldm[k] * M(GMC_LE_LAT, r) * M.rcp(GMC_AREA_CELL, r);
    if (r + 1 <= g.nlat - 2) s += -ldm[k + p] * M(GMC_LE_LAT, r) * M.rcp(GMC_AREA_CELL, r + 1);
    if (r == 0) s += A.pole[2] * M(GMC_LE_LAT, 0) / (double)g.nlon / M(GMC_AREA_CELL, 0);
    if (r == g.nhalf - 1) s += -A.pole[3] * M(GMC_LE_LAT, r) / (double)g.nlon / M(GMC_AREA_CELL, g.nlat - 1);
    double spt = 0.0;
    if (o.coriolis == 1) {
      const double pt = T.pv_lat[k], mf = T.mfy[k];
      if (r + 1 <= g.nlat - 2) {
        const double w1 = M(GMC_FULL_TANGENT_WGT1, r + 1) * 0.5;
        const double qa = lqv[k + p], qb = lqv[(r + 1) * p + im];
        s += w1 * (qa * (T.pv_lon[k + p] + pt) + qb * (T.pv_lon[(r + 1) * p + im] + pt));
        spt += w1 * (qa + qb) * mf;
      }
      if (r >= 1) {
        const double w2 = M(GMC_FULL_TANGENT_WGT2, r) * 0.5;
        const double qa = lqv[k], qb = lqv[r * p + im];
        s += w2 * (qa * (T.pv_lon[k] + pt) + qb * (T.pv_lon[r * p + im] + pt));
        spt += w2 * (qa + qb) * mf;
      }
      spt += lqhu(r, i) * 0.5 *
             (M(GMC_HALF_TANGENT_WGT1, r) * (T.mfx[r * p + im] + T.mfx[k]) +
              M(GMC_HALF_TANGENT_WGT2, r) * (T.mfx[(r + 1) * p + im] + T.mfx[(r + 1) * p + i]));
    } else {
      spt += lqhu(r, i) * T.mfx_lat[k];
    }
    A.w[W_MFY][k] = s;
    A.w[W_PVLAT][k] = spt;
    // u_lat = mfx_lat / m_lat transposed likewise; on the vertex rows next to
    // the poles with Stokes pole vorticity a_mf adds the pole term
    const double lul = spt * T.dlat[0][k];
    double lmf = lul / T.m_lat[k];
    if (o.coriolis != 1) lmf += ldv[k] * -T.pv_lat[k];
    A.w[W_ULAT][k] = lul;
    A.w[W_MFXLAT][k] = lmf;
  }
}

// interp_pv transposed onto pv (vertex rows), then pv = (vor + f) / m_vtx
// transposed: W_VOR = lam vor, W_MVTX = lam m_vtx.  The upwind weights df of
// every edge come from the tape.
__global__ void a_pv(Grid g, MeshR M, Opt o, Tape T, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, R = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || R >= g.nhalf) return;
  const int64_t p = g.pitch, k = R * p + i;
  const double *lpl = A.w[W_PVLON], *lpt = A.w[W_PVLAT];
  double s = 0.0;
  // pv_lon(r, i) reads pv(r-2+kk, i)
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
    const int r = R + 2 - kk;
    if (r < 1 || r > g.nlat - 2) continue;
    const int64_t q = r * p + i;
    s += lpl[q] * T.dlon[2 + kk][q];
  }
  // mirrored reads across the poles: pv_lon(1, i') reads pv(0, i'+half) as f1,
  // pv_lon(nlat-2, i') reads pv(nhalf-1, i'+half) as f4
  if (R == 0 || R == g.nhalf - 1) {
    const int r = R == 0 ? 1 : g.nlat - 2, kk = R == 0 ? 0 : 3;
    const int ii = wrap(i + g.half, g.nlon);   // i' = i - half = i + half (mod nlon)
    const int64_t q = r * p + ii;
    s += lpl[q] * T.dlon[2 + kk][q];
  }
  // pv_lat(R, i') reads pv(R, i'-2+kk)
#pragma unroll
  for (int kk = 0; kk < 4; kk++) {
    const int64_t q = R * p + wrap(i + 2 - kk, g.nlon);
    s += lpt[q] * T.dlat[2 + kk][q];
  }
  const double rv = 1.0 / T.m_vtx[k];
  A.w[W_VOR][k] = s * rv;
  A.w[W_MVTX][k] = -s * T.pv[k] * rv;
}

// Mass fluxes transposed: lam u_f, lam v_f of the star state (from the
// mass fluxes, ke, vor and the upwind weights) and W_MLON, W_MLAT = lam
// m_lon, lam m_lat, from the adjoints a_cori leaves (W_MFX, W_MFY, W_MFXLAT,
// W_MFYLON, W_ULAT, W_VLON, W_PVLON, W_PVLAT).  With Stokes pole vorticity
// the vor of the vertex rows next to the poles is the pole circulation, so
// lam u_lat there gains lam vor summed over that row (W_VOR of a_pv) times
// le_lat / (nlon area_cell): added here, where those rows are read.  The
// point's loads are issued ahead of the pole sums.
__global__ void __launch_bounds__(BX * BY) a_mf(Grid g, MeshR M, Opt o, Tape T, ADPtrs A) {
  __shared__ double s_vor[2];
  const int C0 = blockIdx.x * BX, R0 = blockIdx.y * BY;
  const int64_t p = g.pitch;
  const int i = C0 + threadIdx.x, r = R0 + threadIdx.y;
  const bool pt = i < g.nlon && r < g.nlat;
  const int64_t k = r * p + i;
  const int im = wrap(i - 1, g.nlon), ip = wrap(i + 1, g.nlon);
  const double *lke = A.w[W_KE], *lvo = A.w[W_VOR], *mxl = A.w[W_MFXLAT], *myl = A.w[W_MFYLON];
  auto stokes_row = [&](int rr) { return o.stokes && (rr == 0 || rr == g.nhalf - 1); };
  // the U points' and the V points' outputs in workgroups of their own
  const bool u_pt = pt && blockIdx.z == 0 && r >= 1 && r <= g.nlat - 2;
  const bool v_pt = pt && blockIdx.z == 1 && r <= g.nhalf - 1;
  double u = 0.0, luf = 0.0, v = 0.0, lvf = 0.0, lmfx = 0.0, lmfy = 0.0;
  if (u_pt) {
    u = T.u_f[k];
    // ke(r, i+1) reads u as west, ke(r, i) as east
    const double rac = M.rcp(GMC_AREA_CELL, r);
    luf = lke[r * p + ip] * M(GMC_AREA_LON_WEST, r) * 2 * u * rac + lke[k] * M(GMC_AREA_LON_EAST, r) * 2 * u * rac;
    // vor(r, i) reads u * de_lon(r), vor(r-1, i) reads -u * de_lon(r)
    if (r <= g.nhalf - 1 && !stokes_row(r)) luf += lvo[k] * M(GMC_DE_LON, r) * M.rcp(GMC_AREA_VTX, r);
    if (!stokes_row(r - 1)) luf += -lvo[k - p] * M(GMC_DE_LON, r) * M.rcp(GMC_AREA_VTX, r - 1);
    // the upwind weight of pv_lon reads u_f as its normal component
    luf += A.w[W_PVLON][k] * T.dlon[1][k];
    // lam mfx total: grad_mf + Coriolis (a_cori) + mfx_lat transposed
    lmfx = A.w[W_MFX][k];
    if (r <= g.nhalf - 1) lmfx += M(GMC_HALF_TANGENT_WGT1, r) * (mxl[r * p + ip] + mxl[k]);
    lmfx += M(GMC_HALF_TANGENT_WGT2, r - 1) * (mxl[(r - 1) * p + ip] + mxl[(r - 1) * p + i]);
  }
  if (v_pt) {
    v = T.v_f[k];
    if (r + 1 <= g.nlat - 2) lvf += lke[k + p] * M(GMC_AREA_LAT_NORTH, r) * 2 * v * M.rcp(GMC_AREA_CELL, r + 1);
    if (r >= 1) lvf += lke[k] * M(GMC_AREA_LAT_SOUTH, r) * 2 * v * M.rcp(GMC_AREA_CELL, r);
    if (r == 0) lvf += A.pole[0] * 2 * v / (double)g.nlon;
    if (r == g.nhalf - 1) lvf += A.pole[1] * 2 * v / (double)g.nlon;
    if (!stokes_row(r)) lvf += (lvo[r * p + im] - lvo[k]) * M(GMC_DE_LAT, r) * M.rcp(GMC_AREA_VTX, r);
    lvf += A.w[W_PVLAT][k] * T.dlat[1][k];
    lmfy = A.w[W_MFY][k];
    if (r + 1 <= g.nlat - 2) lmfy += M(GMC_FULL_TANGENT_WGT1, r + 1) * (myl[k + p] + myl[(r + 1) * p + im]);
    if (r >= 1) lmfy += M(GMC_FULL_TANGENT_WGT2, r) * (myl[k] + myl[r * p + im]);
  }
  // the pole terms of lam u_lat: V row 0 is read by rows 0 and 1, V row
  // nhalf-1 by row nhalf-1 (one wavefront sums each row)
  const bool north_rows = R0 <= g.nhalf - 1 && g.nhalf - 1 < R0 + BY;
  if (o.stokes && (R0 == 0 || north_rows)) {
    if (threadIdx.y == 0 && R0 == 0) {
      const double sv = wave_zonal_sum_batched(g.nlon, threadIdx.x, lvo);
      if (threadIdx.x == 0) s_vor[0] = -sv * M(GMC_LE_LAT, 0) / (double)g.nlon / M(GMC_AREA_CELL, 0);
    }
    if (threadIdx.y == 1 && north_rows) {
      const int vr = g.nhalf - 1;
      const double sv = wave_zonal_sum_batched(g.nlon, threadIdx.x, lvo + vr * p);
      if (threadIdx.x == 0) s_vor[1] = sv * M(GMC_LE_LAT, vr) / (double)g.nlon / M(GMC_AREA_CELL, g.nlat - 1);
    }
    __syncthreads();   // the same for the whole workgroup
    // lam u_lat of V row vr gains d; lam mfx_lat gains d / m_lat
    if (u_pt && r == 1)
      lmfx += M(GMC_HALF_TANGENT_WGT2, 0) * (s_vor[0] / T.m_lat[ip] + s_vor[0] / T.m_lat[i]);
    if (u_pt && r == g.nhalf - 1)
      lmfx += M(GMC_HALF_TANGENT_WGT1, r) * (s_vor[1] / T.m_lat[r * p + ip] + s_vor[1] / T.m_lat[k]);
  }
  if (u_pt) {
    A.s_uf[k] += lmfx * T.m_lon[k] + luf;
    A.w[W_MLON][k] = lmfx * u - A.w[W_VLON][k] * T.v_lon[k] / T.m_lon[k];
  }
  if (v_pt) {
    double lul = A.w[W_ULAT][k];
    if (o.stokes && r == 0) lul += s_vor[0];
    if (o.stokes && r == g.nhalf - 1) lul += s_vor[1];
    A.s_vf[k] += lmfy * T.m_lat[k] + lvf;
    A.w[W_MLAT][k] = lmfy * v - lul * T.u_lat[k] / T.m_lat[k];
  }
}

// m_lon, m_lat, m_vtx and m = (gz - gzs) / g transposed onto the star gz
__global__ void a_m(Grid g, MeshR M, Opt o, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, R = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || R >= g.nlat) return;
  A.s_gz[R * g.pitch + i] += am_point(g, M, o.grav, A.w[W_MLON], A.w[W_MLAT], A.w[W_MVTX], R, i);
}

__global__ void a_add(Grid g, int rows, const double *a, const double *b, double *out) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= rows) return;
  const int64_t k = r * g.pitch + i;
  out[k] = a[k] + b[k];
}

// zero / copy up to six whole fields (fld doubles each), field blockIdx.y:
// fill and copy kernels rather than memset / memcpy nodes, so that the
// adjoint step can be captured in a graph
struct FieldSet {
  double *p[6];
};
__global__ void __launch_bounds__(256) k_fill_fields(FieldSet f, int64_t fld) {
  double *p = f.p[blockIdx.y];
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < fld; k += (int64_t)gridDim.x * 256) p[k] = 0.0;
}
__global__ void __launch_bounds__(256) k_copy_fields(FieldSet d, FieldSet src, int64_t fld) {
  double *q = d.p[blockIdx.y];
  const double *p = src.p[blockIdx.y];
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < fld; k += (int64_t)gridDim.x * 256) q[k] = p[k];
}

// cost term and its gradient on the observed cells: J_o = 1/2 sum (gz - y)^2
// / sigma^2; lam gz = (gz - y) / sigma^2; the per-observation terms go to jo
__global__ void k_obs(Grid g, int nobs, const int64_t *__restrict__ cell, const double *__restrict__ y,
                      double inv_var, const double *__restrict__ gz, double *lam_gz, double *jo) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nobs) return;
  const int64_t c = cell[k];
  const int r = (int)(c / g.nlon), i = (int)(c % g.nlon);
  const double d = gz[r * g.pitch + i] - y[k];
  jo[k] = 0.5 * d * d * inv_var;
  lam_gz[r * g.pitch + i] = d * inv_var;   // cells are distinct
}

}  // namespace
}  // namespace gmc

using namespace gmc;

struct gmc_tlad {
  gmc_ctx *c = nullptr;
  double *slab = nullptr;
  size_t fld = 0;                    // doubles per field
  double *X[4][X_N] = {};            // perturbation / adjoint of state levels 1..3
  double *W[W_N] = {};
  double *pole = nullptr;
  size_t lds_uv = 0, lds_gz = 0;   // dynamic LDS of a_uvrow, a_gzrow
  RowItem *d_items = nullptr;      // work items of a_uvrow, then of a_gzrow
  int n_uv_items = 0, n_gz_items = 0;
  // tapes of two checkpoint segments (one being written while the other is
  // read): tapes[(b tape_steps + s) 3 + st] of stage st of step s in buffer b
  double *tape_slab = nullptr;
  int tape_steps = 0;
  std::vector<TapeOut> tapes;
  // the adjoint sweep of a segment runs on astream while the next segment's
  // taping forward runs on the context's stream
  hipStream_t astream = nullptr;
  hipEvent_t ev_fwd = nullptr, ev_adj[2] = {}, ev_join = nullptr, ev_ck0 = nullptr;
  // gmc_gradient's host side: pinned copies of three fields (x0, then the
  // gradient) and the background term's gradient
  HostStage hbuf, obuf;
  std::vector<double> gb;
  std::vector<double *> ck;          // saved states (checkpoints), kNState fields each
  double *obs_buf = nullptr;         // observation cells, values and cost terms
  int obs_cap = 0;
  // graphs of a segment's taping forward (kind 0), of its adjoint sweep
  // (kind 1) and of a plain segment of the window forward (kind 2), keyed by
  // (kind, levels, steps, tape buffer, dt)
  struct SegGraph { int kind, o, n, steps, buf; double dt; hipGraphExec_t exec; };
  std::vector<SegGraph> graphs;
  std::string err;
};

namespace {

int terr(gmc_tlad *a, int code, const std::string &m) {
  if (a) a->err = m;
  return code;
}
int thip(gmc_tlad *a, hipError_t e, const char *what) {
  if (e == hipSuccess) return GMC_OK;
  return terr(a, e == hipErrorOutOfMemory ? GMC_ENOMEM : GMC_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

Opt opt_of(gmc_ctx *c) {
  return Opt{c->cfg.g, c->cfg.pv_scheme, c->cfg.upwind_order_pv, c->cfg.coriolis_scheme, c->cfg.pv_pole_stokes,
             c->cfg.upwind_wgt_pv};
}

Tape tape_of(gmc_ctx *c, int s) {
  double **S = c->state[s];
  return Tape{S[S_M_LON], S[S_M_LAT], S[S_M_VTX], S[S_MFX_LON], S[S_MFY_LAT], S[S_MFX_LAT], S[S_MFY_LON],
              S[S_U_LAT], S[S_V_LON], S[S_PV], S[S_PV_LON], S[S_PV_LAT], S[S_U_F], S[S_V_F]};
}

// stage levels and dt of stage st of the context's scheme (as fused_integrate)
void stage_of(int scheme, int o, int n, double dt, int st, int &so, int &ss, int &sn, double &h) {
  so = o;
  if (scheme == GMC_TIME_EULER) { ss = o; sn = n; h = dt; return; }
  if (st == 0) { ss = o; sn = n; h = scheme == GMC_TIME_WRFRK3 ? dt / 3.0 : dt / 2.0; }
  else if (st == 1) { ss = n; sn = 3; h = dt / 2.0; }
  else { ss = 3; sn = n; h = dt; }
}

int nstages(gmc_ctx *c) { return c->cfg.time_scheme == GMC_TIME_EULER ? 1 : 3; }

// TL of one stage around the tape just computed by ops_stage_tape
void tl_stage(gmc_tlad *a, int o, int s, int n, double h) {
  gmc_ctx *c = a->c;
  const Grid &g = c->grid;
  const Mesh M{c->d_mesh, g.nlat};
  const Opt op = opt_of(c);
  const Tape T = tape_of(c, s);
  const dim3 b(BX, BY);
  TLPtrs P{a->X[o][X_U], a->X[o][X_V], a->X[o][X_GZ], a->X[s][X_GZ], a->X[s][X_UF], a->X[s][X_VF], {},
           a->X[n][X_U], a->X[n][X_V], a->X[n][X_GZ], a->pole};
  for (int k = 0; k < W_N; k++) P.w[k] = a->W[k];
  hipLaunchKernelGGL(t_m, grid2d(g, g.nlat), b, 0, c->stream, g, M, op, P);
  hipLaunchKernelGGL(t_mf, grid2d(g, g.nlat), b, 0, c->stream, g, M, T, P);
  hipLaunchKernelGGL(t_pole, dim3(2), dim3(kWave), 0, c->stream, g, M, T, P);
  hipLaunchKernelGGL(t_ke_pv, grid2d(g, g.nlat), b, 0, c->stream, g, M, op, T, P);
  hipLaunchKernelGGL(t_interp, grid2d(g, g.nlat), b, 0, c->stream, g, op, T, P);
  hipLaunchKernelGGL(t_tend, grid2d(g, g.nlat), b, 0, c->stream, g, M, op, T, P);
  hipLaunchKernelGGL(t_gznew, grid2d(g, g.nlat), b, 0, c->stream, g, h, P);
  ops_filter(c, 0, 0, a->W[W_GZNEW], a->W[W_GZF]);
  ops_filter(c, 1, 0, a->W[W_GZNEW], a->X[n][X_GZ]);
  hipLaunchKernelGGL(t_uv, grid2d(g, g.nlat), b, 0, c->stream, g, M, h, P);
  ops_filter(c, 0, 0, a->X[n][X_U], a->X[n][X_UF]);
  ops_filter(c, 0, 1, a->X[n][X_V], a->X[n][X_VF]);
}

// AD of one stage around its tape T (ops_stage_tape): consumes the adjoint
// of the new state (level n, cleared as it is read) and adds into levels o
// and s.  The two row kernels of the update and its filters, then the
// tendencies' transposes from the Coriolis term back to m.  The last of
// them, a_m, adds into the star gz adjoint, which the next stage reads as
// its new state's: with emit_am false it is left pending for the next
// stage's a_gzrow (pending true there).
void ad_stage(gmc_tlad *a, int o, int s, int n, double h, const Tape &T, bool pending, bool emit_am) {
  gmc_ctx *c = a->c;
  const Grid &g = c->grid;
  const MeshR M{c->d_mesh, c->d_mrows, g.nlat};
  const Opt op = opt_of(c);
  const dim3 b(BX, BY);
  double **Xo = a->X[o], **Xs = a->X[s], **Xn = a->X[n], **W = a->W;
  ADPtrs A{Xn[X_U], Xn[X_V], Xn[X_GZ], Xn[X_UF], Xn[X_VF], Xo[X_U], Xo[X_V], Xo[X_GZ], Xs[X_GZ], Xs[X_UF], Xs[X_VF],
           {}, a->pole};
  for (int k = 0; k < W_N; k++) A.w[k] = W[k];
  UVRow U{Xn[X_U], Xn[X_V], Xn[X_UF], Xn[X_VF], Xo[X_U], Xo[X_V], W[W_DU], W[W_DV]};
  hipLaunchKernelGGL(a_uvrow, dim3(a->n_uv_items), dim3(NR), a->lds_uv, c->stream, g, c->big, h, a->d_items, U);
  GZRow Z{W[W_DU], W[W_DV], Xn[X_GZ], Xo[X_GZ], W[W_KE], W[W_DGZ], a->pole, W[W_MLON], W[W_MLAT], W[W_MVTX],
          pending ? 1 : 0};
  hipLaunchKernelGGL(a_gzrow, dim3(a->n_gz_items), dim3(NR), a->lds_gz, c->stream, g, M, c->big, c->small1, h,
                     op.grav, a->d_items + a->n_uv_items, Z);
  const dim3 g2 = grid2d(g, g.nlat);
  hipLaunchKernelGGL(a_cori, dim3(g2.x, g2.y, 2), b, 0, c->stream, g, M, op, T, A);
  hipLaunchKernelGGL(a_pv, grid2d(g, g.nhalf), b, 0, c->stream, g, M, op, T, A);
  hipLaunchKernelGGL(a_mf, dim3(g2.x, g2.y, 2), b, 0, c->stream, g, M, op, T, A);
  if (emit_am) hipLaunchKernelGGL(a_m, grid2d(g, g.nlat), b, 0, c->stream, g, M, op, A);
}

dim3 fields_grid(gmc_tlad *a, int nf) {
  return dim3((unsigned)std::min<int64_t>(((int64_t)a->fld + 255) / 256, 512), (unsigned)nf);
}
int fill_fields(gmc_tlad *a, double *const *p, int nf) {
  FieldSet f{};
  for (int k = 0; k < nf; k++) f.p[k] = p[k];
  hipLaunchKernelGGL(k_fill_fields, fields_grid(a, nf), dim3(256), 0, a->c->stream, f, (int64_t)a->fld);
  return thip(a, hipGetLastError(), "zero");
}
int copy_fields(gmc_tlad *a, double *const *dst, double *const *src, int nf) {
  FieldSet d{}, q{};
  for (int k = 0; k < nf; k++) { d.p[k] = dst[k]; q.p[k] = src[k]; }
  hipLaunchKernelGGL(k_copy_fields, fields_grid(a, nf), dim3(256), 0, a->c->stream, d, q, (int64_t)a->fld);
  return thip(a, hipGetLastError(), "copy");
}

int zero_level(gmc_tlad *a, int lv) { return fill_fields(a, a->X[lv], X_N); }

// the state fields one step reads and writes (u, v, gz, u_f, v_f, gz_f)
constexpr int kStateFields[] = {S_U_LON, S_V_LAT, S_GZ, S_U_F, S_V_F, S_GZ_F};
constexpr int kNState = 6;

int copy_state(gmc_tlad *a, double *const *dst, double *const *src) { return copy_fields(a, dst, src, kNState); }

// snapshot of level lv into buf[6] (or back)
int save_level(gmc_tlad *a, int lv, double *const *buf) {
  double *src[kNState];
  for (int f = 0; f < kNState; f++) src[f] = a->c->state[lv][kStateFields[f]];
  return copy_state(a, buf, src);
}
int load_level(gmc_tlad *a, int lv, double *const *buf) {
  double *dst[kNState];
  for (int f = 0; f < kNState; f++) dst[f] = a->c->state[lv][kStateFields[f]];
  return copy_state(a, dst, buf);
}

// TL of one step: the nonlinear state advances o -> n (ops stages, the same
// bits as gmc_step), the perturbation X[o] -> X[n]
int tl_step(gmc_tlad *a, int o, int n, double dt) {
  gmc_ctx *c = a->c;
  for (int st = 0; st < nstages(c); st++) {
    int so, ss, sn;
    double h;
    stage_of(c->cfg.time_scheme, o, n, dt, st, so, ss, sn, h);
    int rc = ops_stage_tape(c, so, ss, sn, h);
    if (rc) return rc;
    tl_stage(a, so, ss, sn, h);
  }
  return thip(a, hipGetLastError(), "tl step");
}

// Work lists of the row kernels: each u row, v group and gz row is cut
// into column chunks of about kAdChunkWork / taps columns (the LDS traffic of
// a workgroup grows with taps x columns); the pole rows, whose gz items form
// the pole sums, stay whole.  Heaviest items first.
constexpr int kAdChunkWork = 16384;

void add_chunks(std::vector<RowItem> &v, int kind, int r0, int r1, int taps, int nlon, bool whole) {
  const int w = whole ? nlon : std::min(nlon, std::max(64, (kAdChunkWork / std::max(taps, 1) + 63) / 64 * 64));
  const int parts = (nlon + w - 1) / w;
  for (int k = 0; k < parts; k++)
    v.push_back(RowItem{kind, r0, r1, (int)((int64_t)nlon * k / parts), (int)((int64_t)nlon * (k + 1) / parts), 0});
}

int build_row_items(gmc_tlad *a) {
  gmc_ctx *c = a->c;
  const Grid &g = c->grid;
  std::vector<int32_t> bl(g.nlat), bt(g.nhalf), sl(g.nlat);
  int rc = thip(a, hipMemcpy(bl.data(), c->big.ngrid_lon, sizeof(int32_t) * g.nlat, hipMemcpyDeviceToHost), "taps");
  if (!rc) rc = thip(a, hipMemcpy(bt.data(), c->big.ngrid_lat, sizeof(int32_t) * g.nhalf, hipMemcpyDeviceToHost), "taps");
  if (!rc) rc = thip(a, hipMemcpy(sl.data(), c->small1.ngrid_lon, sizeof(int32_t) * g.nlat, hipMemcpyDeviceToHost), "taps");
  if (rc) return rc;
  auto taps = [](int n) { return std::max(n, 1); };
  std::vector<RowItem> uv, gz;
  for (int r = 0; r < g.nlat; r++) add_chunks(uv, RI_U, r, r, taps(bl[r]), g.nlon, r == 0 || r == g.nlat - 1);
  for (int q = 0; q < v_groups(g.nhalf); q++) {
    int ra, rb, t = 0;
    v_group(q, g.nhalf, ra, rb);
    for (int r = ra; r <= rb; r++) t += taps(bt[r]);
    add_chunks(uv, RI_V, ra, rb, t, g.nlon, false);
  }
  for (int r = 0; r < g.nlat; r++)
    add_chunks(gz, RI_GZ, r, r, taps(bl[r]) + taps(sl[r]), g.nlon, r == 0 || r == g.nlat - 1);
  auto work = [&](const RowItem &it) {
    int t = 0;
    if (it.kind == RI_U) t = taps(bl[it.r0]);
    else if (it.kind == RI_GZ) t = taps(bl[it.r0]) + taps(sl[it.r0]);
    else
      for (int r = it.r0; r <= it.r1; r++) t += taps(bt[r]);
    return (int64_t)t * (it.c1 - it.c0);
  };
  auto heavy_first = [&](const RowItem &x, const RowItem &y) { return work(x) > work(y); };
  std::stable_sort(uv.begin(), uv.end(), heavy_first);
  std::stable_sort(gz.begin(), gz.end(), heavy_first);
  const size_t mb = c->big.max_taps, ms = c->small1.max_taps;
  for (const RowItem &it : uv) {
    const size_t w = it.c1 - it.c0, G = it.r1 - it.r0 + 1;
    a->lds_uv = std::max(a->lds_uv, sizeof(double) * (it.kind == RI_U ? mb + w + mb : G * (mb + w + mb + w)));
  }
  for (const RowItem &it : gz) {
    const size_t w = it.c1 - it.c0;
    a->lds_gz = std::max(a->lds_gz, sizeof(double) * (mb + ms + (w + mb) + (w + ms) + w));
  }
  if (std::max(a->lds_uv, a->lds_gz) > 65536)
    return terr(a, GMC_EINVAL, "rows too long for the adjoint's row kernels (64 KiB of LDS)");
  a->n_uv_items = (int)uv.size();
  a->n_gz_items = (int)gz.size();
  uv.insert(uv.end(), gz.begin(), gz.end());
  rc = thip(a, hipMalloc((void **)&a->d_items, sizeof(RowItem) * uv.size()), "items");
  if (!rc) rc = thip(a, hipMemcpy(a->d_items, uv.data(), sizeof(RowItem) * uv.size(), hipMemcpyHostToDevice), "items");
  return rc;
}

}  // namespace

extern "C" {

int gmc_tlad_create(gmc_ctx *c, gmc_tlad **out) {
  if (!out) return GMC_EINVAL;
  *out = nullptr;
  if (!c || !c->prepared) return GMC_ESTATE;
  // no tangent linear of divergence damping or of the tvd limiter
  if (c->cfg.use_div_damp || c->cfg.pv_scheme == GMC_PV_TVD) return GMC_EINVAL;
  gmc_tlad *a = new gmc_tlad();
  a->c = c;
  const Grid &g = c->grid;
  a->fld = (size_t)g.nlat * g.pitch;
  const size_t n = a->fld * (3 * X_N + W_N) + 8;
  int rc = build_row_items(a);
  if (!rc) rc = thip(a, hipMalloc((void **)&a->slab, n * sizeof(double)), "tlad alloc");
  if (!rc) rc = thip(a, hipMemset(a->slab, 0, n * sizeof(double)), "tlad memset");
  if (!rc) rc = thip(a, hipStreamCreateWithFlags(&a->astream, hipStreamNonBlocking), "stream");
  for (hipEvent_t *e : {&a->ev_fwd, &a->ev_adj[0], &a->ev_adj[1], &a->ev_join, &a->ev_ck0})
    if (!rc) rc = thip(a, hipEventCreateWithFlags(e, hipEventDisableTiming), "event");
  if (rc) {
    gmc_tlad_destroy(a);
    return rc;
  }
  double *p = a->slab;
  for (int lv = 1; lv <= 3; lv++)
    for (int f = 0; f < X_N; f++) { a->X[lv][f] = p; p += a->fld; }
  for (int k = 0; k < W_N; k++) { a->W[k] = p; p += a->fld; }
  a->pole = p;
  *out = a;
  return GMC_OK;
}

int gmc_tlad_destroy(gmc_tlad *a) {
  if (!a) return GMC_EINVAL;
  if (a->c) (void)hipStreamSynchronize(a->c->stream);
  if (a->astream) (void)hipStreamSynchronize(a->astream);
  for (auto &sg : a->graphs) (void)hipGraphExecDestroy(sg.exec);
  for (hipEvent_t e : {a->ev_fwd, a->ev_adj[0], a->ev_adj[1], a->ev_join, a->ev_ck0})
    if (e) (void)hipEventDestroy(e);
  stage_free(a->hbuf);
  stage_free(a->obuf);
  if (a->astream) (void)hipStreamDestroy(a->astream);
  if (a->slab) (void)hipFree(a->slab);
  if (a->d_items) (void)hipFree(a->d_items);
  if (a->tape_slab) (void)hipFree(a->tape_slab);
  for (double *p : a->ck) (void)hipFree(p);
  if (a->obs_buf) (void)hipFree(a->obs_buf);
  delete a;
  return GMC_OK;
}

const char *gmc_tlad_last_error(const gmc_tlad *a) { return a ? a->err.c_str() : "null object"; }

static int var_ptr(gmc_tlad *a, int32_t level, int32_t var, double **p, int *rows) {
  if (level < 1 || level > 3 || var < 0 || var > 4) return terr(a, GMC_EINVAL, "bad level or variable");
  *p = a->X[level][var];
  *rows = (var == GMC_TLAD_V || var == GMC_TLAD_VF) ? a->c->grid.nhalf : a->c->grid.nlat;
  return GMC_OK;
}

int gmc_tlad_upload(gmc_tlad *a, int32_t level, int32_t var, const double *host, int64_t ld) {
  if (!a || !host) return GMC_EINVAL;
  double *p;
  int rows;
  int rc = var_ptr(a, level, var, &p, &rows);
  if (rc) return rc;
  const Grid &g = a->c->grid;
  if (ld < g.nlon) return terr(a, GMC_EINVAL, "ld < nlon");
  rc = thip(a, stage_copy(a->c->stage, a->c->stream, true, p, g.pitch, const_cast<double *>(host), ld, g.nlon, rows),
            "upload");
  return rc;
}

int gmc_tlad_download(gmc_tlad *a, int32_t level, int32_t var, double *host, int64_t ld) {
  if (!a || !host) return GMC_EINVAL;
  double *p;
  int rows;
  int rc = var_ptr(a, level, var, &p, &rows);
  if (rc) return rc;
  const Grid &g = a->c->grid;
  if (ld < g.nlon) return terr(a, GMC_EINVAL, "ld < nlon");
  rc = thip(a, stage_copy(a->c->stage, a->c->stream, false, p, g.pitch, host, ld, g.nlon, rows), "download");
  return rc;
}

int gmc_tlad_zero(gmc_tlad *a) {
  if (!a) return GMC_EINVAL;
  int rc = GMC_OK;
  for (int lv = 1; lv <= 3 && !rc; lv++) rc = zero_level(a, lv);
  if (!rc) rc = thip(a, hipStreamSynchronize(a->c->stream), "sync");
  return rc;
}

int gmc_tl_step(gmc_tlad *a, double dt, int32_t nsteps) {
  if (!a) return GMC_EINVAL;
  if (nsteps < 0 || !(dt > 0)) return terr(a, GMC_EINVAL, "bad step arguments");
  gmc_ctx *c = a->c;
  derived_forget(c);
  int rc = GMC_OK;
  for (int s = 0; s < nsteps && !rc; s++) {
    const int o = c->old_level, n = c->new_level;
    // u_f = BF(u), v_f = BF(v) of the step's initial state (gmc_prepare or
    // the previous step's update) in tangent-linear form
    ops_filter(c, 0, 0, a->X[o][X_U], a->X[o][X_UF]);
    ops_filter(c, 0, 1, a->X[o][X_V], a->X[o][X_VF]);
    rc = tl_step(a, o, n, dt);
    if (!rc) std::swap(c->old_level, c->new_level);
  }
  if (!rc) rc = thip(a, hipStreamSynchronize(c->stream), "sync");
  return rc;
}

namespace {

// room for `nstates` saved states (kNState fields each) in a->ck, kept
// from call to call
int alloc_traj(gmc_tlad *a, int nstates) {
  const size_t want = (size_t)nstates * kNState;
  while (a->ck.size() < want) {
    double *p = nullptr;
    int rc = thip(a, hipMalloc((void **)&p, a->fld * sizeof(double)), "trajectory alloc");
    if (rc) return rc;
    a->ck.push_back(p);
  }
  return GMC_OK;
}

// device copies of the observation cells and values, and their cost terms
int alloc_obs(gmc_tlad *a, int nobs) {
  if (nobs <= a->obs_cap) return GMC_OK;
  int rc = thip(a, hipStreamSynchronize(a->c->stream), "sync");
  if (a->obs_buf) (void)hipFree(a->obs_buf);
  a->obs_buf = nullptr;
  a->obs_cap = 0;
  if (!rc) rc = thip(a, hipMalloc((void **)&a->obs_buf, sizeof(double) * 3 * (size_t)nobs), "obs alloc");
  if (!rc) a->obs_cap = nobs;
  return rc;
}


constexpr int kTapeFields = 26;   // the fields of a TapeOut
constexpr int kAdSegment = 8;     // checkpoint interval of gmc_ad_step

// room for the tapes of `steps` steps; graphs captured over an older store
// are dropped with it
int ensure_tapes(gmc_tlad *a, int steps) {
  if (steps <= a->tape_steps) return GMC_OK;
  int rc = thip(a, hipStreamSynchronize(a->c->stream), "sync");
  if (!rc) rc = thip(a, hipStreamSynchronize(a->astream), "sync");
  if (rc) return rc;
  for (auto &sg : a->graphs) (void)hipGraphExecDestroy(sg.exec);
  a->graphs.clear();
  if (a->tape_slab) (void)hipFree(a->tape_slab);
  a->tape_slab = nullptr;
  a->tape_steps = 0;
  const size_t n = 2 * (size_t)steps * 3 * kTapeFields * a->fld;
  rc = thip(a, hipMalloc((void **)&a->tape_slab, n * sizeof(double)), "tape alloc");
  // the taping stages write each field on the rows where the forward forms
  // it; the adjoint also reads mfx on the pole rows, where the forward's
  // flux is zero, so the rows never written must hold zeros
  if (!rc) rc = thip(a, hipMemset(a->tape_slab, 0, n * sizeof(double)), "tape clear");
  if (rc) return rc;
  a->tape_steps = steps;
  a->tapes.assign(2 * (size_t)steps * 3, TapeOut{});
  double *p = a->tape_slab;
  for (TapeOut &t : a->tapes) {
    double **f[kTapeFields] = {&t.m_lon,   &t.m_lat,   &t.m_vtx,   &t.mfx,     &t.mfy,     &t.mfx_lat, &t.mfy_lon,
                               &t.u_lat,   &t.v_lon,   &t.pv,      &t.pv_lon,  &t.pv_lat,  &t.u_f,     &t.v_f,
                               &t.dlon[0], &t.dlon[1], &t.dlon[2], &t.dlon[3], &t.dlon[4], &t.dlon[5], &t.dlat[0],
                               &t.dlat[1], &t.dlat[2], &t.dlat[3], &t.dlat[4], &t.dlat[5]};
    for (int k = 0; k < kTapeFields; k++) { *f[k] = p; p += a->fld; }
    // the tangential mass fluxes are read by the adjoint of Coriolis scheme 2 only
    if (a->c->cfg.coriolis_scheme == 1) t.mfx_lat = t.mfy_lon = nullptr;
  }
  return GMC_OK;
}

TapeOut *tapes_of(gmc_tlad *a, int buf) { return a->tapes.data() + (size_t)buf * a->tape_steps * 3; }

Tape tape_from(const TapeOut &t) {
  Tape T{t.m_lon, t.m_lat, t.m_vtx, t.mfx, t.mfy, t.mfx_lat, t.mfy_lon, t.u_lat, t.v_lon, t.pv, t.pv_lon,
         t.pv_lat, t.u_f, t.v_f, {}, {}};
  for (int j = 0; j < 6; j++) {
    T.dlon[j] = t.dlon[j];
    T.dlat[j] = t.dlat[j];
  }
  return T;
}

// replay on `stream` the graph of `body` for this key, capturing it (on the
// context's stream) on first use
int graph_run(gmc_tlad *a, int kind, int o, int n, int steps, int buf, double dt, hipStream_t stream,
              const std::function<int()> &body) {
  gmc_ctx *c = a->c;
  hipGraphExec_t exec = nullptr;
  for (auto &sg : a->graphs)
    if (sg.kind == kind && sg.o == o && sg.n == n && sg.steps == steps && sg.buf == buf && sg.dt == dt)
      exec = sg.exec;
  if (!exec) {
    hipGraph_t graph = nullptr;
    int rc = thip(a, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    const int rc2 = body();
    rc = thip(a, hipStreamEndCapture(c->stream, &graph), "end capture");
    if (rc2) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc2;
    }
    if (rc) return rc;
    rc = thip(a, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    (void)hipGraphDestroy(graph);
    if (rc) return rc;
    a->graphs.push_back({kind, o, n, steps, buf, dt, exec});
  }
  return thip(a, hipGraphLaunch(exec, stream), "graph launch");
}

// The segment's steps forward from level `start` (alternating with level
// `other`) through the taping fused stages (the same bits as gmc_step);
// stage st of step s leaves its tape in tape buffer buf.
int seg_forward_body(gmc_tlad *a, int start, int other, double dt, int steps, int buf) {
  gmc_ctx *c = a->c;
  TapeOut *tp = tapes_of(a, buf);
  int cur = start, rc = GMC_OK;
  for (int s = 0; s < steps && !rc; s++) {
    const int nxt = cur == start ? other : start;
    for (int st = 0; st < nstages(c) && !rc; st++) {
      int so, ss, sn;
      double h;
      stage_of(c->cfg.time_scheme, cur, nxt, dt, st, so, ss, sn, h);
      rc = fused_stage_tape(c, so, ss, sn, h, &tp[(size_t)s * 3 + st]);
    }
    cur = nxt;
  }
  return rc;
}

// Adjoint of the segment's steps, last first, around their tapes.  The
// adjoint of the segment's end state is in X[xin] and the other of levels
// (o, n) and level 3 are zero; each step consumes its input level and leaves
// the adjoint of its initial state in the other, so the levels alternate.
int seg_adjoint_body(gmc_tlad *a, int o, int n, double dt, int xin, int steps, int buf) {
  gmc_ctx *c = a->c;
  const TapeOut *tp = tapes_of(a, buf);
  int x_in = xin, x_out = xin == o ? n : o;
  // a stage's star level is the next stage's new level (in this reverse
  // order, across steps too), so a_m is deferred to the next stage's a_gzrow
  // except after the segment's last stage
  bool pending = false;
  for (int s = steps - 1; s >= 0; s--) {
    for (int st = nstages(c) - 1; st >= 0; st--) {
      int so, ss, sn;
      double h;
      stage_of(c->cfg.time_scheme, x_out, x_in, dt, st, so, ss, sn, h);
      const bool last = s == 0 && st == 0;
      ad_stage(a, so, ss, sn, h, tape_from(tp[(size_t)s * 3 + st]), pending, last);
      pending = !last;
    }
    std::swap(x_in, x_out);
  }
  return thip(a, hipGetLastError(), "adjoint segment");
}

// Forward over the window from level o, keeping the state every `every`
// steps in ck; the last segment runs through the taping stages, its tapes
// left in buffer (segments - 1) % 2.  cur is the level of the end state.
// ck0_saved, when given, is recorded once the initial state is saved.
int window_forward(gmc_tlad *a, int o, int n, double dt, int nsteps, int every, double *const *ck, int &cur,
                   hipEvent_t ck0_saved = nullptr) {
  gmc_ctx *c = a->c;
  int rc = ensure_tapes(a, std::min(every, nsteps));
  const int nck = (nsteps + every - 1) / every;
  cur = o;
  for (int q = 0; q < nck && !rc; q++) {
    const int start = cur, other = cur == o ? n : o, steps = std::min(every, nsteps - q * every);
    rc = save_level(a, cur, &ck[(size_t)q * kNState]);
    if (!rc && q == 0 && ck0_saved) rc = thip(a, hipEventRecord(ck0_saved, c->stream), "event");
    if (rc) break;
    if (q < nck - 1) {
      // a plain segment: the fused steps as one graph
      rc = graph_run(a, 2, start, other, steps, -1, dt, c->stream, [&] {
        int cc = start, r2 = GMC_OK;
        for (int s = 0; s < steps && !r2; s++) {
          const int nxt = cc == start ? other : start;
          r2 = fused_integrate(c, cc, nxt, dt);
          cc = nxt;
        }
        return r2;
      });
    } else {
      const int buf = (nck - 1) % 2;
      rc = graph_run(a, 0, start, other, steps, buf, dt, c->stream,
                     [&] { return seg_forward_body(a, start, other, dt, steps, buf); });
    }
    if (steps % 2) cur = other;
  }
  return rc;
}

// Reverse sweep over the window's checkpoint segments, last first, after
// window_forward.  The adjoint of segment q replays from tape buffer q % 2
// on the adjoint stream while segment q-1 runs forward from its checkpoint
// on the context's stream into the other buffer (events order each buffer's
// writer after its last reader).  On entry the adjoint of the window's end
// state is in X[o] and X[n], X[3] are zero; on return the adjoint of its
// initial state is in X[o] (u_f, v_f parts still separate), X[n] and X[3]
// are zero and the state is back at the window's start in level o.
int window_adjoint(gmc_tlad *a, int o, int n, double dt, int nsteps, int every, double *const *ck) {
  hipStream_t F = a->c->stream, A = a->astream;
  const int nck = (nsteps + every - 1) / every;
  bool read[2] = {false, false};
  int xin = o;
  int rc = thip(a, hipEventRecord(a->ev_fwd, F), "event");
  for (int q = nck - 1; q >= 0 && !rc; q--) {
    const int steps = std::min(every, nsteps - q * every), buf = q % 2, xout = xin == o ? n : o, x0 = xin;
    rc = thip(a, hipStreamWaitEvent(A, a->ev_fwd, 0), "event wait");
    if (!rc) rc = graph_run(a, 1, x0, xout, steps, buf, dt, A,
                            [&] { return seg_adjoint_body(a, o, n, dt, x0, steps, buf); });
    if (!rc) rc = thip(a, hipEventRecord(a->ev_adj[buf], A), "event");
    read[buf] = true;
    if (steps % 2) xin = xout;
    if (rc || q == 0) break;
    const int pbuf = (q - 1) % 2;
    if (read[pbuf]) rc = thip(a, hipStreamWaitEvent(F, a->ev_adj[pbuf], 0), "event wait");
    if (!rc) rc = load_level(a, o, &ck[(size_t)(q - 1) * kNState]);
    if (!rc) rc = graph_run(a, 0, o, n, every, pbuf, dt, F,
                            [&] { return seg_forward_body(a, o, n, dt, every, pbuf); });
    if (!rc) rc = thip(a, hipEventRecord(a->ev_fwd, F), "event");
  }
  if (!rc) rc = thip(a, hipEventRecord(a->ev_join, A), "event");
  if (!rc) rc = thip(a, hipStreamWaitEvent(F, a->ev_join, 0), "event wait");
  if (!rc && xin != o) {
    rc = copy_fields(a, a->X[o], a->X[xin], X_N);
    if (!rc) rc = zero_level(a, xin);
  }
  if (!rc) rc = load_level(a, o, &ck[0]);
  return rc;
}

// lam u += BF(lam u_f), lam v += BF(lam v_f) for the window's initial state,
// whose u_f, v_f are the filters of its u, v
int fold_filters(gmc_tlad *a, int o) {
  gmc_ctx *c = a->c;
  const Grid &g = c->grid;
  ops_filter(c, 0, 0, a->X[o][X_UF], a->W[W_T1]);
  ops_filter(c, 0, 1, a->X[o][X_VF], a->W[W_T2]);
  const dim3 b(BX, BY);
  hipLaunchKernelGGL(a_add, grid2d(g, g.nlat), b, 0, c->stream, g, g.nlat, a->W[W_T1], a->X[o][X_U], a->X[o][X_U]);
  hipLaunchKernelGGL(a_add, grid2d(g, g.nhalf), b, 0, c->stream, g, g.nhalf, a->W[W_T2], a->X[o][X_V], a->X[o][X_V]);
  double *uvf[2] = {a->X[o][X_UF], a->X[o][X_VF]};
  return fill_fields(a, uvf, 2);
}

}  // namespace

int gmc_ad_step(gmc_tlad *a, double dt, int32_t nsteps) {
  if (!a) return GMC_EINVAL;
  if (nsteps < 0 || !(dt > 0)) return terr(a, GMC_EINVAL, "bad step arguments");
  gmc_ctx *c = a->c;
  derived_forget(c);
  const int o = c->old_level, n = c->new_level;
  int rc = fused_prepare(c);
  // the adjoint enters on u, v, gz only
  double *uvf[2] = {a->X[o][X_UF], a->X[o][X_VF]};
  if (!rc) rc = fill_fields(a, uvf, 2);
  if (!rc) rc = zero_level(a, n);
  if (!rc) rc = zero_level(a, 3);
  if (!rc && nsteps > 0) {
    const int every = std::min<int>(nsteps, kAdSegment);
    int cur = o;
    rc = alloc_traj(a, (nsteps + every - 1) / every);
    if (!rc) rc = window_forward(a, o, n, dt, nsteps, every, a->ck.data(), cur);
    if (!rc) rc = window_adjoint(a, o, n, dt, nsteps, every, a->ck.data());
  }
  if (!rc) rc = fold_filters(a, o);
  if (!rc) rc = thip(a, hipStreamSynchronize(c->stream), "sync");
  return rc;
}

int gmc_gradient(gmc_tlad *a, double dt, int32_t nsteps, int32_t ckpt_every, int32_t nobs, const int64_t *obs_cell,
                 const double *obs_value, double sigma_o, const double *ub, const double *vb, const double *gzb,
                 double sigma_b, int64_t ld, double *cost, double *grad_u, double *grad_v, double *grad_gz) {
  if (!a) return GMC_EINVAL;
  gmc_ctx *c = a->c;
  derived_forget(c);
  const Grid &g = c->grid;
  if (nsteps < 1 || ckpt_every < 1 || nobs < 0 || (nobs && (!obs_cell || !obs_value)) || !(sigma_o > 0) ||
      !(sigma_b > 0) || !ub || !vb || !gzb || ld < g.nlon || !cost || !grad_u || !grad_v || !grad_gz || !(dt > 0))
    return terr(a, GMC_EINVAL, "bad gradient arguments");
  {
    std::vector<uint8_t> seen((size_t)g.nlat * g.nlon, 0);
    for (int k = 0; k < nobs; k++) {
      if (obs_cell[k] < 0 || obs_cell[k] >= (int64_t)g.nlat * g.nlon)
        return terr(a, GMC_EINVAL, "observation cell out of range");
      if (seen[obs_cell[k]]++) return terr(a, GMC_EINVAL, "observation cells must be distinct");
    }
  }
  const int o = c->old_level, n = c->new_level;
  const int nck = (nsteps + ckpt_every - 1) / ckpt_every;
  int rc = fused_prepare(c);
  if (!rc) rc = alloc_traj(a, nck);
  if (!rc) rc = alloc_obs(a, nobs);
  double *const *ck = a->ck.data();
  // forward with checkpoints every ckpt_every steps
  int cur = o;
  if (!rc) rc = window_forward(a, o, n, dt, nsteps, ckpt_every, ck, cur, a->ev_ck0);
  // background term J_b = 1/2 |x0 - xb|^2 / sigma_b^2 over u (rows between
  // the poles), v and gz, and its gradient, on the host while the window
  // runs: x0 comes from its checkpoint copy on the adjoint stream
  const size_t fsz = (size_t)g.nlat * g.nlon;
  const int rows[3] = {g.nlat, g.nhalf, g.nlat}, r0[3] = {1, 0, 0}, r1[3] = {g.nlat - 1, g.nhalf, g.nlat};
  const double *xb[3] = {ub, vb, gzb};
  double *gr[3] = {grad_u, grad_v, grad_gz};
  const double ib = 1.0 / (sigma_b * sigma_b);
  double jb = 0.0;
  if (!rc && a->hbuf.cap < 3 * fsz) {
    stage_free(a->hbuf);
    rc = thip(a, hipHostMalloc((void **)&a->hbuf.p, 3 * fsz * sizeof(double), hipHostMallocDefault), "pinned alloc");
    if (rc) a->hbuf.p = nullptr;
    else a->hbuf.cap = 3 * fsz;
  }
  if (!rc) rc = thip(a, hipStreamWaitEvent(a->astream, a->ev_ck0, 0), "event wait");
  for (int f = 0; f < 3 && !rc; f++)
    rc = thip(a, hipMemcpy2DAsync(a->hbuf.p + f * fsz, sizeof(double) * g.nlon, ck[f], sizeof(double) * g.pitch,
                                  sizeof(double) * g.nlon, rows[f], hipMemcpyDeviceToHost, a->astream), "x0 copy");
  for (int lv = 1; lv <= 3 && !rc; lv++) {
    FieldSet f{};
    for (int k = 0; k < X_N; k++) f.p[k] = a->X[lv][k];
    hipLaunchKernelGGL(k_fill_fields, fields_grid(a, X_N), dim3(256), 0, a->astream, f, (int64_t)a->fld);
  }
  if (!rc) rc = thip(a, hipEventRecord(a->ev_join, a->astream), "event");
  if (!rc) rc = thip(a, hipStreamSynchronize(a->astream), "sync");
  if (!rc) {
    a->gb.assign(3 * fsz, 0.0);
    for (int f = 0; f < 3; f++)
      for (int r = r0[f]; r < r1[f]; r++)
        for (int i = 0; i < g.nlon; i++) {
          const size_t k = (size_t)r * g.nlon + i;
          const double d = a->hbuf.p[f * fsz + k] - xb[f][(size_t)r * ld + i];
          jb += 0.5 * d * d * ib;
          a->gb[f * fsz + k] = d * ib;
        }
  }
  // observation term at the final state and its adjoint
  int64_t *d_cell = (int64_t *)a->obs_buf;
  double *d_y = a->obs_buf + nobs, *d_jo = a->obs_buf + 2 * (size_t)nobs;
  double jo = 0.0;
  if (!rc) rc = thip(a, hipStreamWaitEvent(c->stream, a->ev_join, 0), "event wait");   // levels cleared
  if (!rc && nobs && a->obuf.cap < 3 * (size_t)nobs) {
    stage_free(a->obuf);
    rc = thip(a, hipHostMalloc((void **)&a->obuf.p, 3 * (size_t)nobs * sizeof(double), hipHostMallocDefault),
              "pinned alloc");
    if (rc) a->obuf.p = nullptr;
    else a->obuf.cap = 3 * (size_t)nobs;
  }
  if (!rc && nobs) {
    std::memcpy(a->obuf.p, obs_cell, sizeof(int64_t) * nobs);
    std::memcpy(a->obuf.p + nobs, obs_value, sizeof(double) * nobs);
    rc = thip(a, hipMemcpyAsync(d_cell, a->obuf.p, sizeof(double) * 2 * nobs, hipMemcpyHostToDevice, c->stream),
              "obs copy");
    if (!rc) {
      hipLaunchKernelGGL(k_obs, dim3((nobs + 255) / 256), dim3(256), 0, c->stream, g, nobs, d_cell, d_y,
                         1.0 / (sigma_o * sigma_o), c->state[cur][S_GZ], a->X[o][X_GZ], d_jo);
      rc = thip(a, hipMemcpyAsync(a->obuf.p + 2 * (size_t)nobs, d_jo, sizeof(double) * nobs, hipMemcpyDeviceToHost,
                                  c->stream), "obs copy");
      // waiting here, at the window's end, before the adjoint sweep is
      // queued measured 0.3 ms faster than queueing the sweep behind k_obs
      if (!rc) rc = thip(a, hipStreamSynchronize(c->stream), "sync");
      for (int k = 0; k < nobs && !rc; k++) jo += a->obuf.p[2 * (size_t)nobs + k];
    }
  }
  // reverse over the checkpoint segments, each recomputed forward once
  if (!rc) rc = window_adjoint(a, o, n, dt, nsteps, ckpt_every, ck);
  if (!rc) rc = fold_filters(a, o);
  // the gradient: the adjoint of x0 plus the background term's
  const int xv[3] = {X_U, X_V, X_GZ};
  for (int f = 0; f < 3 && !rc; f++)
    rc = thip(a, hipMemcpy2DAsync(a->hbuf.p + f * fsz, sizeof(double) * g.nlon, a->X[o][xv[f]],
                                  sizeof(double) * g.pitch, sizeof(double) * g.nlon, rows[f], hipMemcpyDeviceToHost,
                                  c->stream), "gradient copy");
  if (!rc) rc = thip(a, hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  for (int f = 0; f < 3; f++)
    for (int r = 0; r < rows[f]; r++) {
      double *__restrict__ out = gr[f] + (size_t)r * ld;
      if (r < r0[f] || r >= r1[f]) {
        std::memset(out, 0, sizeof(double) * g.nlon);
        continue;
      }
      const double *__restrict__ x = a->hbuf.p + f * fsz + (size_t)r * g.nlon;
      const double *__restrict__ xg = a->gb.data() + f * fsz + (size_t)r * g.nlon;
      for (int i = 0; i < g.nlon; i++) out[i] = x[i] + xg[i];
    }
  *cost = jo + jb;
  return zero_level(a, o);
}

}  // extern "C"
