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

// the forward trajectory of one stage (fields of the star state, its tend)
struct Tape {
  const double *m_lon, *m_lat, *m_vtx, *mfx, *mfy, *mfx_lat, *mfy_lon, *u_lat, *v_lon, *pv, *pv_lon, *pv_lat;
  const double *u_f, *v_f;
};

struct Opt {
  double grav;
  int pv_scheme, order, coriolis, stokes;
  double wgt;
};

__device__ __forceinline__ double dsign(double t) { return copysign(1.0, t); }

// derivatives of pv_upwind_edge (gmc_point.hpp) with respect to t, n and the
// four pv values, on the branch the trajectory took
struct UpwD {
  double dt, dn, df[4];
};
__device__ __forceinline__ UpwD upwind_d(const Opt &o, double t, double n, int lat_edge, double f1, double f2,
                                         double f3, double f4) {
  UpwD d;
  if (o.pv_scheme == GMC_PV_MIDPOINT) {
    d.dt = d.dn = 0.0;
    d.df[0] = d.df[3] = 0.0;
    d.df[1] = d.df[2] = 0.5;
    return d;
  }
  const double S = sqrt(lat_edge ? t * t + n * n : n * n + t * t);
  const double den = S + kEps;
  const double b = fabs(t) / den;
  const double dir = copysign(1.0, t);
  const double w = o.wgt;
  double up, u1, u2, u3, u4;
  if (o.order == 1) {
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
  const double dbdt = dsign(t) / den - q * t, dbdn = -q * n;
  d.dt = dbdt * (up - c);
  d.dn = dbdn * (up - c);
  d.df[0] = b * u1;
  d.df[1] = b * u2 + (1 - b) * 0.5;
  d.df[2] = b * u3 + (1 - b) * 0.5;
  d.df[3] = b * u4;
  return d;
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
  const double *n_u, *n_v, *n_gz, *n_uf, *n_vf;
  double *o_u, *o_v, *o_gz, *s_gz, *s_uf, *s_vf;
  double *w[W_N];
  double *pole;  // [0..1] ke, [2..3] dmfdlat, [4..5] vor: sums of pole-row adjoints
};

// W_T1 = BF(lam u_f) + lam u, W_T2 = BF(lam v_f) + lam v on entry (the
// update's filter transposed); this forms the smoothing transpose and the
// pgf transpose: W_DU = h lam u*, W_DV = h lam v*, lam_o u/v, W_GZF = lam gz_f
__global__ void a_uv(Grid g, double h, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  const double *lu = A.w[W_T1], *lv = A.w[W_T2];
  if (r >= 1 && r <= g.nlat - 2) {
    A.o_u[k] += lu[k];
    A.w[W_DU][k] = h * lu[k];
  }
  if (r <= g.nhalf - 1) {
    const int last = g.nhalf - 1;
    const double wgt = 0.6;
    double ls = (r == 0 || r == last) ? (1 - wgt) * lv[k] : lv[k];
    if (r == 1) ls += wgt * lv[k - p];
    if (r == last - 1) ls += wgt * lv[k + p];
    A.o_v[k] += ls;
    A.w[W_DV][k] = h * ls;
  }
}

// lam gz_f from the pgf terms: pgf_lon = (gzf(i+1) - gzf(i)) / de_lon with
// lam pgf_lon = -W_DU, pgf_lat likewise
__global__ void a_pgf(Grid g, Mesh M, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  const int im = wrap(i - 1, g.nlon);
  const double *du = A.w[W_DU], *dv = A.w[W_DV];
  double s = 0.0;
  if (r >= 1 && r <= g.nlat - 2) s += (du[k] - du[r * p + im]) / M(GMC_DE_LON, r);
  if (r >= 1) s += -dv[k - p] / M(GMC_DE_LAT, r - 1);
  if (r <= g.nhalf - 1) s += dv[k] / M(GMC_DE_LAT, r);
  A.w[W_GZF][k] = s;
}

// lam gz_new = BF(lam gz_f) + SF(lam gz) arrives in W_T1 + W_T2
__global__ void a_gznew(Grid g, double h, double grav, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t k = r * g.pitch + i;
  const double lg = A.w[W_T1][k] + A.w[W_T2][k];
  A.o_gz[k] += lg;
  A.w[W_DGZ][k] = -(h * lg) * grav;   // lam dmfdlon = lam dmfdlat
}

// lam ke from grad_ke (lam dkedlon = -lam du, lam dkedlat = -lam dv)
__global__ void a_ke(Grid g, Mesh M, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  const int im = wrap(i - 1, g.nlon);
  const double *du = A.w[W_DU], *dv = A.w[W_DV];
  double s = 0.0;
  if (r >= 1 && r <= g.nlat - 2) s += (du[k] - du[r * p + im]) / M(GMC_DE_LON, r);
  if (r >= 1) s += -dv[k - p] / M(GMC_DE_LAT, r - 1);
  if (r <= g.nhalf - 1) s += dv[k] / M(GMC_DE_LAT, r);
  A.w[W_KE][k] = s;
}

// sums over the pole rows: ke (pole cells), dmfdlat (pole cells, from
// W_DGZ), vor (vertex rows next to the poles); one wavefront per pole
__global__ void a_pole(Grid g, int which, ADPtrs A) {
  const int north = blockIdx.x, lane = threadIdx.x;
  const int64_t p = g.pitch;
  if (which == 0) {
    const int cr = north ? g.nlat - 1 : 0;
    const double *ke = A.w[W_KE] + cr * p, *dg = A.w[W_DGZ] + cr * p;
    const double sk = wave_zonal_sum(g.nlon, lane, [&](int i) { return ke[i]; });
    const double sd = wave_zonal_sum(g.nlon, lane, [&](int i) { return dg[i]; });
    if (lane == 0) {
      A.pole[north] = sk;
      A.pole[2 + north] = sd;
    }
  } else {
    const int vr = north ? g.nhalf - 1 : 0;
    const double *vo = A.w[W_VOR] + vr * p;
    const double s = wave_zonal_sum(g.nlon, lane, [&](int i) { return vo[i]; });
    if (lane == 0) A.pole[4 + north] = s;
  }
}

// Coriolis and grad_mf transposed: lam mfx, lam pv_lon (U rows), lam mfy,
// lam pv_lat (V rows); lam qhv = W_DU / h ... carried as lam du (= h lam u*):
// the forward tendencies enter the update as h du, so W_DU already holds
// lam du.
__global__ void a_cori(Grid g, Mesh M, Opt o, Tape T, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  const int im = wrap(i - 1, g.nlon), ip = wrap(i + 1, g.nlon);
  const double *lqv = A.w[W_DU];   // lam qhv = lam du
  const double *ldv = A.w[W_DV];   // lam qhu = -lam dv
  const double *ldm = A.w[W_DGZ];  // lam dmfdlon = lam dmfdlat
  auto lqhu = [&](int rr, int ii) { return -ldv[rr * p + ii]; };
  if (r >= 1 && r <= g.nlat - 2) {
    // lam mfx: grad_mf
    double s = (ldm[k] - ldm[r * p + ip]) * M(GMC_LE_LON, r) / M(GMC_AREA_CELL, r);
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
  }
  if (r <= g.nhalf - 1) {
    // lam mfy: grad_mf (cell rows r and r + 1) and the pole sums
    double s = 0.0;
    if (r >= 1) s += ldm[k] * M(GMC_LE_LAT, r) / M(GMC_AREA_CELL, r);
    if (r + 1 <= g.nlat - 2) s += -ldm[k + p] * M(GMC_LE_LAT, r) / M(GMC_AREA_CELL, r + 1);
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
  }
}

// interp_pv transposed onto pv (vertex rows), then pv = (vor + f) / m_vtx
// transposed: W_VOR = lam vor, W_MVTX = lam m_vtx
__global__ void a_pv(Grid g, Mesh M, Opt o, Tape T, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, R = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || R >= g.nhalf) return;
  const int64_t p = g.pitch, k = R * p + i;
  auto PV = [&](int rr, int ii) { mirror_half(rr, ii, g); return T.pv[rr * p + ii]; };
  const double *lpl = A.w[W_PVLON], *lpt = A.w[W_PVLAT];
  double s = 0.0;
  // pv_lon(r, i) reads pv(r-2+kk, i)
  for (int kk = 0; kk < 4; kk++) {
    const int r = R + 2 - kk;
    if (r < 1 || r > g.nlat - 2) continue;
    const int64_t q = r * p + i;
    const UpwD d = upwind_d(o, T.v_lon[q], T.u_f[q], 0, PV(r - 2, i), PV(r - 1, i), PV(r, i), PV(r + 1, i));
    s += lpl[q] * d.df[kk];
  }
  // mirrored reads across the poles: pv_lon(1, i') reads pv(0, i'+half) as f1,
  // pv_lon(nlat-2, i') reads pv(nhalf-1, i'+half) as f4
  if (R == 0 || R == g.nhalf - 1) {
    const int r = R == 0 ? 1 : g.nlat - 2, kk = R == 0 ? 0 : 3;
    const int ii = wrap(i + g.half, g.nlon);   // i' = i - half = i + half (mod nlon)
    const int64_t q = r * p + ii;
    const UpwD d = upwind_d(o, T.v_lon[q], T.u_f[q], 0, PV(r - 2, ii), PV(r - 1, ii), PV(r, ii), PV(r + 1, ii));
    s += lpl[q] * d.df[kk];
  }
  // pv_lat(R, i') reads pv(R, i'-2+kk)
  for (int kk = 0; kk < 4; kk++) {
    const int ii = wrap(i + 2 - kk, g.nlon);
    const int64_t q = R * p + ii;
    const UpwD d = upwind_d(o, T.u_lat[q], T.v_f[q], 1, T.pv[R * p + wrap(ii - 2, g.nlon)],
                            T.pv[R * p + wrap(ii - 1, g.nlon)], T.pv[q], T.pv[R * p + wrap(ii + 1, g.nlon)]);
    s += lpt[q] * d.df[kk];
  }
  A.w[W_PV][k] = s;
  A.w[W_VOR][k] = s / T.m_vtx[k];
  A.w[W_MVTX][k] = -s * T.pv[k] / T.m_vtx[k];
}

// u_lat = mfx_lat / m_lat and v_lon = mfy_lon / m_lon transposed, with the
// upwind weights' dependence on t: W_MFXLAT, W_MFYLON = lam of the tangential
// mass fluxes, W_ULAT = lam u_lat, W_VLON = lam v_lon (kept for a_mf)
__global__ void a_lat(Grid g, Mesh M, Opt o, Tape T, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  auto PV = [&](int rr, int ii) { mirror_half(rr, ii, g); return T.pv[rr * p + ii]; };
  if (r <= g.nhalf - 1) {
    const UpwD d = upwind_d(o, T.u_lat[k], T.v_f[k], 1, T.pv[r * p + wrap(i - 2, g.nlon)],
                            T.pv[r * p + wrap(i - 1, g.nlon)], T.pv[k], T.pv[r * p + wrap(i + 1, g.nlon)]);
    double lul = A.w[W_PVLAT][k] * d.dt;
    if (o.stokes && r == 0) lul += -A.pole[4] * M(GMC_LE_LAT, 0) / (double)g.nlon / M(GMC_AREA_CELL, 0);
    if (o.stokes && r == g.nhalf - 1)
      lul += A.pole[5] * M(GMC_LE_LAT, r) / (double)g.nlon / M(GMC_AREA_CELL, g.nlat - 1);
    double lmf = lul / T.m_lat[k];
    if (o.coriolis != 1) lmf += -A.w[W_DV][k] * T.pv_lat[k];
    A.w[W_ULAT][k] = lul;
    A.w[W_MFXLAT][k] = lmf;
  }
  if (r >= 1 && r <= g.nlat - 2) {
    const UpwD d = upwind_d(o, T.v_lon[k], T.u_f[k], 0, PV(r - 2, i), PV(r - 1, i), PV(r, i), PV(r + 1, i));
    const double lvl = A.w[W_PVLON][k] * d.dt;
    double lmf = lvl / T.m_lon[k];
    if (o.coriolis != 1) lmf += A.w[W_DU][k] * T.pv_lon[k];
    A.w[W_VLON][k] = lvl;
    A.w[W_MFYLON][k] = lmf;
  }
}

// mass fluxes transposed: lam u_f, lam v_f of the star state (from the mass
// fluxes, ke, vor and the upwind weights) and lam m_lon, lam m_lat
__global__ void a_mf(Grid g, Mesh M, Opt o, Tape T, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, k = r * p + i;
  const int im = wrap(i - 1, g.nlon), ip = wrap(i + 1, g.nlon);
  auto PV = [&](int rr, int ii) { mirror_half(rr, ii, g); return T.pv[rr * p + ii]; };
  const double *lke = A.w[W_KE], *lvo = A.w[W_VOR];
  auto stokes_row = [&](int rr) { return o.stokes && (rr == 0 || rr == g.nhalf - 1); };
  if (r >= 1 && r <= g.nlat - 2) {
    // lam mfx total: grad_mf + Coriolis (a_cori) + mfx_lat transposed
    double lmfx = A.w[W_MFX][k];
    if (r <= g.nhalf - 1)
      lmfx += M(GMC_HALF_TANGENT_WGT1, r) * (A.w[W_MFXLAT][r * p + ip] + A.w[W_MFXLAT][k]);
    if (r - 1 >= 0)
      lmfx += M(GMC_HALF_TANGENT_WGT2, r - 1) * (A.w[W_MFXLAT][(r - 1) * p + ip] + A.w[W_MFXLAT][(r - 1) * p + i]);
    const double u = T.u_f[k];
    double luf = lmfx * T.m_lon[k];
    // ke(r, i+1) reads u as west, ke(r, i) as east
    luf += lke[r * p + ip] * M(GMC_AREA_LON_WEST, r) * 2 * u / M(GMC_AREA_CELL, r) +
           lke[k] * M(GMC_AREA_LON_EAST, r) * 2 * u / M(GMC_AREA_CELL, r);
    // vor(r, i) reads u * de_lon(r), vor(r-1, i) reads -u * de_lon(r)
    if (r <= g.nhalf - 1 && !stokes_row(r)) luf += lvo[k] * M(GMC_DE_LON, r) / M(GMC_AREA_VTX, r);
    if (r - 1 >= 0 && !stokes_row(r - 1)) luf += -lvo[k - p] * M(GMC_DE_LON, r) / M(GMC_AREA_VTX, r - 1);
    // the upwind weight of pv_lon reads u_f as its normal component
    const UpwD d = upwind_d(o, T.v_lon[k], u, 0, PV(r - 2, i), PV(r - 1, i), PV(r, i), PV(r + 1, i));
    luf += A.w[W_PVLON][k] * d.dn;
    A.s_uf[k] += luf;
    A.w[W_MLON][k] = lmfx * u - A.w[W_VLON][k] * T.v_lon[k] / T.m_lon[k];
  }
  if (r <= g.nhalf - 1) {
    double lmfy = A.w[W_MFY][k];
    if (r + 1 <= g.nlat - 2)
      lmfy += M(GMC_FULL_TANGENT_WGT1, r + 1) * (A.w[W_MFYLON][k + p] + A.w[W_MFYLON][(r + 1) * p + im]);
    if (r >= 1) lmfy += M(GMC_FULL_TANGENT_WGT2, r) * (A.w[W_MFYLON][k] + A.w[W_MFYLON][r * p + im]);
    const double v = T.v_f[k];
    double lvf = lmfy * T.m_lat[k];
    if (r + 1 <= g.nlat - 2) lvf += lke[k + p] * M(GMC_AREA_LAT_NORTH, r) * 2 * v / M(GMC_AREA_CELL, r + 1);
    if (r >= 1) lvf += lke[k] * M(GMC_AREA_LAT_SOUTH, r) * 2 * v / M(GMC_AREA_CELL, r);
    if (r == 0) lvf += A.pole[0] * 2 * v / (double)g.nlon;
    if (r == g.nhalf - 1) lvf += A.pole[1] * 2 * v / (double)g.nlon;
    if (!stokes_row(r)) lvf += (lvo[r * p + im] - lvo[k]) * M(GMC_DE_LAT, r) / M(GMC_AREA_VTX, r);
    const UpwD d = upwind_d(o, T.u_lat[k], v, 1, T.pv[r * p + wrap(i - 2, g.nlon)], T.pv[r * p + im], T.pv[k],
                            T.pv[r * p + ip]);
    lvf += A.w[W_PVLAT][k] * d.dn;
    A.s_vf[k] += lvf;
    A.w[W_MLAT][k] = lmfy * v - A.w[W_ULAT][k] * T.u_lat[k] / T.m_lat[k];
  }
}

// m_lon, m_lat, m_vtx and m = (gz - gzs) / g transposed onto the star gz
__global__ void a_m(Grid g, Mesh M, Opt o, ADPtrs A) {
  const int i = blockIdx.x * BX + threadIdx.x, R = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || R >= g.nlat) return;
  const int64_t p = g.pitch, k = R * p + i;
  const int im = wrap(i - 1, g.nlon);
  const double *ll = A.w[W_MLON], *lt = A.w[W_MLAT], *lv = A.w[W_MVTX];
  double s = 0.0;
  if (R >= 1 && R <= g.nlat - 2) s += 0.5 * (ll[k] + ll[R * p + im]);
  if (R <= g.nhalf - 1) {
    s += 0.5 * lt[k];
    s += (lv[k] + lv[R * p + im]) * M(GMC_AREA_SUBCELL2, R) / M(GMC_AREA_VTX, R);
  }
  if (R >= 1) {
    s += 0.5 * lt[k - p];
    s += (lv[k - p] + lv[(R - 1) * p + im]) * M(GMC_AREA_SUBCELL1, R) / M(GMC_AREA_VTX, R - 1);
  }
  A.s_gz[k] += s / o.grav;
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
  double *tape1[5] = {};   // m_lon, m_lat, m_vtx, u_f, v_f of stage 1's star state
  // graphs of one adjoint step, keyed by (old, new, dt)
  struct StepGraph { int o, n; double dt; hipGraphExec_t exec; };
  std::vector<StepGraph> graphs;
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
// of the new state (level n) and adds into levels o and s
void ad_stage(gmc_tlad *a, int o, int s, int n, double h, const Tape &T) {
  gmc_ctx *c = a->c;
  const Grid &g = c->grid;
  const Mesh M{c->d_mesh, g.nlat};
  const Opt op = opt_of(c);
  const dim3 b(BX, BY);
  ADPtrs A{a->X[n][X_U], a->X[n][X_V], a->X[n][X_GZ], a->X[n][X_UF], a->X[n][X_VF], a->X[o][X_U], a->X[o][X_V],
           a->X[o][X_GZ], a->X[s][X_GZ], a->X[s][X_UF], a->X[s][X_VF], {}, a->pole};
  for (int k = 0; k < W_N; k++) A.w[k] = a->W[k];
  // u_f = BF(u), v_f = BF(v) of the new state: lam u += BF(lam u_f)
  ops_filter(c, 0, 0, A.n_uf, a->W[W_T1]);
  ops_filter(c, 0, 1, A.n_vf, a->W[W_T2]);
  hipLaunchKernelGGL(a_add, grid2d(g, g.nlat), b, 0, c->stream, g, g.nlat, a->W[W_T1], A.n_u, a->W[W_T1]);
  hipLaunchKernelGGL(a_add, grid2d(g, g.nhalf), b, 0, c->stream, g, g.nhalf, a->W[W_T2], A.n_v, a->W[W_T2]);
  hipLaunchKernelGGL(a_uv, grid2d(g, g.nlat), b, 0, c->stream, g, h, A);
  hipLaunchKernelGGL(a_pgf, grid2d(g, g.nlat), b, 0, c->stream, g, M, A);
  // gz_f = BF(gz_new), gz = SF(gz_new)
  ops_filter(c, 0, 0, a->W[W_GZF], a->W[W_T1]);
  ops_filter(c, 1, 0, A.n_gz, a->W[W_T2]);
  hipLaunchKernelGGL(a_gznew, grid2d(g, g.nlat), b, 0, c->stream, g, h, op.grav, A);
  hipLaunchKernelGGL(a_ke, grid2d(g, g.nlat), b, 0, c->stream, g, M, A);
  hipLaunchKernelGGL(a_pole, dim3(2), dim3(kWave), 0, c->stream, g, 0, A);
  hipLaunchKernelGGL(a_cori, grid2d(g, g.nlat), b, 0, c->stream, g, M, op, T, A);
  hipLaunchKernelGGL(a_pv, grid2d(g, g.nhalf), b, 0, c->stream, g, M, op, T, A);
  hipLaunchKernelGGL(a_pole, dim3(2), dim3(kWave), 0, c->stream, g, 1, A);
  hipLaunchKernelGGL(a_lat, grid2d(g, g.nlat), b, 0, c->stream, g, M, op, T, A);
  hipLaunchKernelGGL(a_mf, grid2d(g, g.nlat), b, 0, c->stream, g, M, op, T, A);
  hipLaunchKernelGGL(a_m, grid2d(g, g.nlat), b, 0, c->stream, g, M, op, A);
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

// AD of one step whose old state is in level o: X[n] (adjoint of the step's
// result) -> added into X[o].  The stages run forward once through the fused
// kernels (the same bits as the operator path), each leaving its tape in its
// star level: stage 0 in o, stage 1 in n, stage 2 in 3.  Stage 2 writes its
// new state into level n, over the u_f, v_f of stage 1's tape, so stage 1's
// tape fields that live in level n are kept aside after stage 1.
int ad_step_body(gmc_tlad *a, int o, int n, double dt) {
  gmc_ctx *c = a->c;
  const int ns = nstages(c);
  int rc = GMC_OK;
  for (int st = 0; st < ns && !rc; st++) {
    int so, ss, sn;
    double h;
    stage_of(c->cfg.time_scheme, o, n, dt, st, so, ss, sn, h);
    rc = fused_stage_tape(c, so, ss, sn, h);
    if (!rc && st == 1 && ns > 2) {
      double *src[5] = {c->state[ss][S_M_LON], c->state[ss][S_M_LAT], c->state[ss][S_M_VTX], c->state[ss][S_U_F],
                        c->state[ss][S_V_F]};
      rc = copy_fields(a, a->tape1, src, 5);
    }
  }
  for (int st = ns - 1; st >= 0 && !rc; st--) {
    int so, ss, sn;
    double h;
    stage_of(c->cfg.time_scheme, o, n, dt, st, so, ss, sn, h);
    Tape T = tape_of(c, ss);
    if (st == 1 && ns > 2) {
      T.m_lon = a->tape1[0];
      T.m_lat = a->tape1[1];
      T.m_vtx = a->tape1[2];
      T.u_f = a->tape1[3];
      T.v_f = a->tape1[4];
    }
    ad_stage(a, so, ss, sn, h, T);
    rc = zero_level(a, sn);   // the new state's adjoint is consumed
  }
  if (rc) return rc;
  return thip(a, hipGetLastError(), "ad step");
}

}  // namespace

extern "C" {

int gmc_tlad_create(gmc_ctx *c, gmc_tlad **out) {
  if (!out) return GMC_EINVAL;
  *out = nullptr;
  if (!c || !c->prepared) return GMC_ESTATE;
  if (c->cfg.use_div_damp) return GMC_EINVAL;
  gmc_tlad *a = new gmc_tlad();
  a->c = c;
  const Grid &g = c->grid;
  a->fld = (size_t)g.nlat * g.pitch;
  const size_t n = a->fld * (3 * X_N + W_N + 5) + 8;
  int rc = thip(a, hipMalloc((void **)&a->slab, n * sizeof(double)), "tlad alloc");
  if (!rc) rc = thip(a, hipMemset(a->slab, 0, n * sizeof(double)), "tlad memset");
  if (rc) {
    gmc_tlad_destroy(a);
    return rc;
  }
  double *p = a->slab;
  for (int lv = 1; lv <= 3; lv++)
    for (int f = 0; f < X_N; f++) { a->X[lv][f] = p; p += a->fld; }
  for (int k = 0; k < W_N; k++) { a->W[k] = p; p += a->fld; }
  for (int k = 0; k < 5; k++) { a->tape1[k] = p; p += a->fld; }
  a->pole = p;
  *out = a;
  return GMC_OK;
}

int gmc_tlad_destroy(gmc_tlad *a) {
  if (!a) return GMC_EINVAL;
  if (a->c) (void)hipStreamSynchronize(a->c->stream);
  for (auto &sg : a->graphs) (void)hipGraphExecDestroy(sg.exec);
  if (a->slab) (void)hipFree(a->slab);
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
  rc = thip(a, hipMemcpy2DAsync(p, g.pitch * sizeof(double), host, ld * sizeof(double), g.nlon * sizeof(double),
                                rows, hipMemcpyHostToDevice, a->c->stream), "upload");
  if (!rc) rc = thip(a, hipStreamSynchronize(a->c->stream), "sync");
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
  rc = thip(a, hipMemcpy2DAsync(host, ld * sizeof(double), p, g.pitch * sizeof(double), g.nlon * sizeof(double),
                                rows, hipMemcpyDeviceToHost, a->c->stream), "download");
  if (!rc) rc = thip(a, hipStreamSynchronize(a->c->stream), "sync");
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

struct Traj {
  std::vector<double *> b;   // steps x kNState fields
  ~Traj() {
    for (auto p : b)
      if (p) (void)hipFree(p);
  }
};

int alloc_traj(gmc_tlad *a, Traj &t, int nstates) {
  t.b.assign((size_t)nstates * kNState, nullptr);
  for (auto &p : t.b) {
    int rc = thip(a, hipMalloc((void **)&p, a->fld * sizeof(double)), "trajectory alloc");
    if (rc) return rc;
  }
  return GMC_OK;
}


// One adjoint step as a graph: the adjoint of the step's result moves from
// X[o] to X[n], X[o] and X[3] are cleared, then ad_step_body.  The step's
// initial state is loaded into level o before the launch.
int ad_step_graph(gmc_tlad *a, int o, int n, double dt) {
  gmc_ctx *c = a->c;
  hipGraphExec_t exec = nullptr;
  for (auto &sg : a->graphs)
    if (sg.o == o && sg.n == n && sg.dt == dt) exec = sg.exec;
  if (!exec) {
    hipGraph_t graph = nullptr;
    int rc = thip(a, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    int rc2 = copy_fields(a, a->X[n], a->X[o], X_N);
    if (!rc2) rc2 = zero_level(a, o);
    if (!rc2) rc2 = zero_level(a, 3);
    if (!rc2) rc2 = ad_step_body(a, o, n, dt);
    rc = thip(a, hipStreamEndCapture(c->stream, &graph), "end capture");
    if (rc2) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc2;
    }
    if (rc) return rc;
    rc = thip(a, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    (void)hipGraphDestroy(graph);
    if (rc) return rc;
    a->graphs.push_back({o, n, dt, exec});
  }
  return thip(a, hipGraphLaunch(exec, c->stream), "graph launch");
}

// adjoint of `steps` steps whose initial states are traj[first ..]: the
// adjoint of the last result is in X[o] on entry, of the first initial state
// on return (u_f, v_f parts still separate)
int ad_sweep(gmc_tlad *a, int o, int n, double dt, Traj &t, int first, int steps) {
  int rc = GMC_OK;
  for (int s = first + steps - 1; s >= first && !rc; s--) {
    rc = load_level(a, o, &t.b[(size_t)s * kNState]);
    if (!rc) rc = ad_step_graph(a, o, n, dt);
  }
  return rc;
}

// one plain forward step (the fused step graph) from level cur into the
// other of levels o, n; cur becomes that level
int fwd_step(gmc_tlad *a, int o, int n, double dt, int &cur) {
  const int nxt = cur == o ? n : o;
  const int rc = fused_step_graph(a->c, cur, nxt, dt);
  cur = nxt;
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
  const int o = c->old_level, n = c->new_level;
  int rc = fused_prepare(c);
  // the adjoint enters on u, v, gz only
  double *uvf[2] = {a->X[o][X_UF], a->X[o][X_VF]};
  if (!rc) rc = fill_fields(a, uvf, 2);
  Traj t;
  if (!rc) rc = alloc_traj(a, t, nsteps + 1);
  // forward: keep the initial state of every step
  int cur = o;
  for (int s = 0; s < nsteps && !rc; s++) {
    rc = save_level(a, cur, &t.b[(size_t)s * kNState]);
    if (!rc) rc = fwd_step(a, o, n, dt, cur);
  }
  if (!rc) rc = ad_sweep(a, o, n, dt, t, 0, nsteps);
  if (!rc) rc = fold_filters(a, o);
  if (!rc && nsteps > 0) rc = load_level(a, o, &t.b[0]);
  if (!rc) rc = thip(a, hipStreamSynchronize(c->stream), "sync");
  return rc;
}

int gmc_gradient(gmc_tlad *a, double dt, int32_t nsteps, int32_t ckpt_every, int32_t nobs, const int64_t *obs_cell,
                 const double *obs_value, double sigma_o, const double *ub, const double *vb, const double *gzb,
                 double sigma_b, int64_t ld, double *cost, double *grad_u, double *grad_v, double *grad_gz) {
  if (!a) return GMC_EINVAL;
  gmc_ctx *c = a->c;
  const Grid &g = c->grid;
  if (nsteps < 1 || ckpt_every < 1 || nobs < 0 || (nobs && (!obs_cell || !obs_value)) || !(sigma_o > 0) ||
      !(sigma_b > 0) || !ub || !vb || !gzb || ld < g.nlon || !cost || !grad_u || !grad_v || !grad_gz || !(dt > 0))
    return terr(a, GMC_EINVAL, "bad gradient arguments");
  for (int k = 0; k < nobs; k++)
    if (obs_cell[k] < 0 || obs_cell[k] >= (int64_t)g.nlat * g.nlon) return terr(a, GMC_EINVAL, "observation cell out of range");
  {
    std::vector<int64_t> sorted(obs_cell, obs_cell + nobs);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      return terr(a, GMC_EINVAL, "observation cells must be distinct");
  }
  const int o = c->old_level, n = c->new_level;
  const int nck = (nsteps + ckpt_every - 1) / ckpt_every;
  int rc = fused_prepare(c);
  Traj ck, seg;
  if (!rc) rc = alloc_traj(a, ck, nck);
  if (!rc) rc = alloc_traj(a, seg, ckpt_every);
  // forward with checkpoints every ckpt_every steps
  int cur = o;
  for (int s = 0; s < nsteps && !rc; s++) {
    if (s % ckpt_every == 0) rc = save_level(a, cur, &ck.b[(size_t)(s / ckpt_every) * kNState]);
    if (!rc) rc = fwd_step(a, o, n, dt, cur);
  }
  // observation term at the final state and its adjoint
  int64_t *d_cell = nullptr;
  double *d_y = nullptr, *d_jo = nullptr;
  double jo = 0.0;
  if (!rc) rc = zero_level(a, o);
  if (!rc) rc = zero_level(a, n);
  if (!rc) rc = zero_level(a, 3);
  if (!rc && nobs) {
    rc = thip(a, hipMalloc((void **)&d_cell, sizeof(int64_t) * nobs), "obs alloc");
    if (!rc) rc = thip(a, hipMalloc((void **)&d_y, sizeof(double) * nobs), "obs alloc");
    if (!rc) rc = thip(a, hipMalloc((void **)&d_jo, sizeof(double) * nobs), "obs alloc");
    if (!rc) rc = thip(a, hipMemcpyAsync(d_cell, obs_cell, sizeof(int64_t) * nobs, hipMemcpyHostToDevice, c->stream), "obs copy");
    if (!rc) rc = thip(a, hipMemcpyAsync(d_y, obs_value, sizeof(double) * nobs, hipMemcpyHostToDevice, c->stream), "obs copy");
    if (!rc) {
      hipLaunchKernelGGL(k_obs, dim3((nobs + 255) / 256), dim3(256), 0, c->stream, g, nobs, d_cell, d_y,
                         1.0 / (sigma_o * sigma_o), c->state[cur][S_GZ], a->X[o][X_GZ], d_jo);
      std::vector<double> h(nobs);
      rc = thip(a, hipMemcpyAsync(h.data(), d_jo, sizeof(double) * nobs, hipMemcpyDeviceToHost, c->stream), "obs copy");
      if (!rc) rc = thip(a, hipStreamSynchronize(c->stream), "sync");
      for (int k = 0; k < nobs; k++) jo += h[k];
    }
  }
  // reverse over the checkpoint segments, recomputing each segment forward
  for (int q = nck - 1; q >= 0 && !rc; q--) {
    const int first = q * ckpt_every, steps = std::min(ckpt_every, nsteps - first);
    // the adjoint of the segment's end state is in X[o]; recompute from the
    // checkpoint without touching it
    rc = load_level(a, o, &ck.b[(size_t)q * kNState]);
    cur = o;
    for (int s = 0; s < steps && !rc; s++) {
      rc = save_level(a, cur, &seg.b[(size_t)s * kNState]);
      if (!rc) rc = fwd_step(a, o, n, dt, cur);
    }
    if (!rc) rc = ad_sweep(a, o, n, dt, seg, 0, steps);
  }
  if (!rc) rc = fold_filters(a, o);
  if (!rc) rc = load_level(a, o, &ck.b[0]);
  if (d_cell) (void)hipFree(d_cell);
  if (d_y) (void)hipFree(d_y);
  if (d_jo) (void)hipFree(d_jo);
  if (rc) return rc;
  // background term on the host: J_b = 1/2 |x0 - xb|^2 / sigma_b^2 over u
  // (rows between the poles), v and gz; its gradient added to the adjoint
  rc = gmc_tlad_download(a, o, GMC_TLAD_U, grad_u, ld);
  if (!rc) rc = gmc_tlad_download(a, o, GMC_TLAD_V, grad_v, ld);
  if (!rc) rc = gmc_tlad_download(a, o, GMC_TLAD_GZ, grad_gz, ld);
  if (rc) return rc;
  std::vector<double> x((size_t)g.nlat * ld);
  const double ib = 1.0 / (sigma_b * sigma_b);
  double jb = 0.0;
  struct F { int field, rows, r0, r1; const double *xb; double *gr; };
  const F fs[3] = {{GMC_U_LON, g.nlat, 1, g.nlat - 1, ub, grad_u}, {GMC_V_LAT, g.nhalf, 0, g.nhalf, vb, grad_v},
                   {GMC_GZ, g.nlat, 0, g.nlat, gzb, grad_gz}};
  for (const F &f : fs) {
    rc = gmc_download(c, o, f.field, x.data(), ld);
    if (rc) return terr(a, rc, "state download");
    for (int r = 0; r < f.rows; r++)
      for (int i = 0; i < g.nlon; i++) {
        const size_t k = (size_t)r * ld + i;
        if (r < f.r0 || r >= f.r1) { f.gr[k] = 0.0; continue; }
        const double d = x[k] - f.xb[k];
        jb += 0.5 * d * d * ib;
        f.gr[k] += d * ib;
      }
  }
  *cost = jo + jb;
  return zero_level(a, o);
}

}  // extern "C"
