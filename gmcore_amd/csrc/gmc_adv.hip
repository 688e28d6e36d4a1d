// gmc_adv.hip -- FFSL tracer advection of one cell-centred advection batch on
// gfx950: the C ABI of include/gmcore_adv.h.
//
// Per step (adv_run, src/dynamics/adv/adv_mod.F90:152-312):
//   k_adv_accum   wind accumulation (adv_batch_mod.F90:218-252, 379-403)
//   k_adv_cfl     RK4 CFL numbers and split divergences (:254-373), only on
//                 the calls whose accumulation window closes
//   k_adv_inner   per cell: inner PPM fluxes of q along x (the lane's U
//                 point; the west face comes from the lane below) and y (the
//                 two V rows either side), then the advective half-steps qx,
//                 qy (ffsl_mod.F90:350-407)
//   k_adv_flux    per cell: outer PPM fluxes of qy along x and qx along y
//                 (ffsl_mod.F90:410), the flux-form update and /m
//                 (adv_mod.F90:183-235) into qp
//   k_adv_fix     per (row, tracer): the positivity fixer and the longitude
//                 damping of the row (adv_mod.F90:236-301), qp -> new level
// The flux kernels read every stencil from global memory (L2): one lane per
// cell, no LDS and no barrier, so many wavefronts hide the latency of the
// data-dependent loads (the upwind cells).  The pole rows, whose y-flux is a
// zonal sum, are one workgroup each.  Only the fixer and the damping, which
// walk a row, keep the row in LDS.
//
// Layout in HBM: dense row-major [rows][pitch] doubles per field, no halo;
// tracer l of level k at q + ((k-1) * ntracers + l) * nlat * pitch.
// Longitude neighbours wrap; rows beyond a pole are the mirrored rows with a
// half-circle shift (the reference's fill_halo on one rank).  Every formula
// keeps the reference's operation order; the build uses -ffp-contract=off.
#include <hip/hip_ext.h>
#include <algorithm>
#include <functional>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gmc_internal.hpp"
#include "gmc_point.hpp"
#include "gmc_hflx.hpp"
#include "gmc_ppm.hpp"
#include "../../include/gmcore_adv.h"

namespace gmc {
namespace {

constexpr int NA = 256;  // threads per workgroup of the row kernels
enum { F_CFLX = 1, F_CFLY = 2, F_NEG = 4 };
enum { KA_ACCUM, KA_CFL, KA_INNER, KA_FLUX, KA_FIX, KA_NCLASS };
const char *const kAdvNames[KA_NCLASS] = {"k_adv_accum", "k_adv_cfl", "k_adv_inner", "k_adv_flux", "k_adv_fix"};

__device__ __forceinline__ int wrap_any(int i, int n) {
  const int r = i % n;
  return r < 0 ? r + n : r;
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

// grid (column chunks of NA, rows): one column per thread; the samples and
// the running sums of the four fields are loaded before any of them is
// stored (with a basis, u_f and mfx_lon share the basis ub, v_f and mfy_lat
// the basis vb)
__global__ void __launch_bounds__(NA) k_adv_accum(AdvGrid g, AccumPtrs p, int mode, int nstep, double scale) {
  const int r = blockIdx.y, i = blockIdx.x * NA + threadIdx.x;
  if (i >= g.nlon) return;
  const int64_t k = r * g.pitch + i;
  const bool vrow = r < g.nhalf;   // the V-point fields have nlat - 1 rows
  double *acc[4] = {p.u, p.v, p.mfx, p.mfy};
  double x[4], a[4] = {0.0, 0.0, 0.0, 0.0};
  if (p.ub) {
    const double xu = p.ub[k] * scale, xv = vrow ? p.vb[k] * scale : 0.0;
    x[0] = xu;
    x[1] = xv;
    x[2] = xu;
    x[3] = xv;
  } else {
    x[0] = p.su[k];
    x[1] = vrow ? p.sv[k] : 0.0;
    x[2] = p.smfx[k];
    x[3] = vrow ? p.smfy[k] : 0.0;
  }
  if (mode != 0) {
#pragma unroll
    for (int f = 0; f < 4; f++)
      if (!(f & 1) || vrow) a[f] = acc[f][k];
  }
#pragma unroll
  for (int f = 0; f < 4; f++) {
    if ((f & 1) && !vrow) continue;
    if (mode == 0) acc[f][k] = x[f];
    else if (mode == 1) acc[f][k] = (a[f] + x[f]) / (double)(nstep + 1);
    else acc[f][k] = a[f] + x[f];
  }
}

// CFL numbers and split divergences (adv_batch_mod.F90:254-373); grid
// (column chunks of NA, row index r): cflx and divx, divy of full row r,
// cfly of half row r, one column per thread; the pole rows' divy from the
// zonal sum of the adjacent V row, by the first chunk's workgroup.
struct CflPtrs {
  const double *u, *v;
  double *cflx, *cfly, *divx, *divy;
  int *flags;
};

__global__ void __launch_bounds__(NA) k_adv_cfl(AdvGrid g, Mesh M, CflPtrs p) {
  const int r = blockIdx.y, n = g.nlon, i = blockIdx.x * NA + threadIdx.x;
  const double R = g.radius, dt = g.dt;
  int bad = 0;
  if (r > 0 && r < g.nlat - 1 && i < n) {
    const double dx = M(GMC_DE_LON, r), cl = M(GMC_FULL_COS_LAT, r);
    const double *u = p.u + r * g.pitch;
    {
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
  if (r < g.nhalf && i < n) {
    const double dy = M(GMC_DE_LAT, r), y0 = R * M(GMC_HALF_LAT, r);
    {
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
  if ((r == 0 || r == g.nlat - 1) && blockIdx.x == 0 && threadIdx.x < kWave) {
    const int h = r == 0 ? 0 : g.nhalf - 1;
    const double sgn = r == 0 ? 1.0 : -1.0;
    const double *vh = p.v + h * g.pitch;
    double pole = wave_zonal_sum(n, threadIdx.x, [&](int k) { return sgn * vh[k]; });
    pole = pole * M(GMC_LE_LAT, h) / n / M(GMC_AREA_CELL, r);
    for (int k = threadIdx.x; k < n; k += kWave) p.divy[r * g.pitch + k] = pole;
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

// Work of the two flux kernels, in blocks of NF = 8 wavefronts.  Items
// [0, npole) are the pole rows (2 per tracer, padded to a multiple of the
// XCD count so that the interior items keep their XCD placement).  An
// interior item covers the cell rows r0 .. r0 + 6 (r0 = 1 + 7 rb) over one
// column segment: wavefront w forms the y-fluxes of V row r0 - 1 + w (into
// LDS, each shared by the cells either side) and, for w < 7, the x-fluxes
// and the update of cell row r0 + w.  Lane l sits on column 63 s - 1 + l;
// lanes 1..63 own the cells 63 s .. 63 s + 62, the x-flux on a cell's west
// face coming from the lane below.  Interior items are numbered (row block,
// segment, tracer), the tracer fastest, and dealt to the XCDs in contiguous
// runs, so the wind rows and the neighbouring tracer rows an item reads are
// already in its XCD's L2.
constexpr int NF = 512;
constexpr int kSeg = kWave - 1;
constexpr int kRowsPerItem = NF / kWave - 1;
struct FluxItem {
  int pole;    // 1: pole row item
  int r, t;    // pole: cell row; interior: first cell row r0; tracer
  int c0;      // first cell of the segment
};
__device__ __forceinline__ FluxItem flux_item(const AdvGrid &g, int ntr) {
  const int npole = (2 * ntr + kNumXcd - 1) / kNumXcd * kNumXcd;
  const int b = blockIdx.x;
  FluxItem it{0, -1, 0, 0};
  if (b < npole) {
    if (b < 2 * ntr) { it.pole = 1; it.r = b % 2 ? g.nlat - 1 : 0; it.t = b / 2; }
    return it;
  }
  const int nseg = (g.nlon + kSeg - 1) / kSeg, nrb = (g.nlat - 2 + kRowsPerItem - 1) / kRowsPerItem;
  const int L = xcd_item(b - npole, nrb * nseg * ntr);
  it.t = L % ntr;
  const int rs = L / ntr, s = rs % nseg, rb = rs / nseg;
  it.r = 1 + kRowsPerItem * rb;
  it.c0 = s * kSeg;
  return it;
}
int flux_blocks(const AdvGrid &g, int ntr) {
  const int npole = (2 * ntr + kNumXcd - 1) / kNumXcd * kNumXcd;
  return npole + ((g.nlat - 2 + kRowsPerItem - 1) / kRowsPerItem) * ((g.nlon + kSeg - 1) / kSeg) * ntr;
}

// Sum over the columns of f(i) in wave_zonal_sum's order (lane l adds f(l),
// f(l + 64), ... then the xor butterfly), by the first wavefront of the
// block; the values are formed four at a time ahead of the additions.
template <class F>
__device__ __forceinline__ double wave_row_sum(int n, F f) {
  const int lane = threadIdx.x;
  double acc = 0.0;
  for (int base = lane; base < n; base += 4 * kWave) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = base + u * kWave < n ? f(base + u * kWave) : 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (base + u * kWave < n) acc = acc + v[u];
  }
  for (int w = 1; w < kWave; w <<= 1) acc = acc + __shfl_xor(acc, w, kWave);
  return acc;
}

// ffsl_calc_tracer_hflx_cell, inner part (ffsl_mod.F90:350-407): the
// advective half-steps qx, qy of every cell
// (NF, 8): eight wavefronts per SIMD (four workgroups per CU); without the
// bound the kernel argument and row registers take more scalar registers
// than that occupancy allows
template <bool VL, int LIM>
__global__ void __launch_bounds__(NF, 8) k_adv_inner(AdvGrid g, Mesh M, int ntr, InnerPtrs p) {
  constexpr int lim = LIM;   // the limiter's branches fold at compile time
  __shared__ double sfy[NF / kWave][kWave];
  __shared__ double bcast;
  const FluxItem it = flux_item(g, ntr);
  const int n = g.nlon;
  const int64_t toff = it.t * p.tstride, P = g.pitch;
  const double *q = p.q + toff;
  const double dt = g.dt;
  if (it.pole) {
    // pole rows: qx = q, qy from the zonal sum of the adjacent V row's flux
    const int r = it.r, h = r == 0 ? 0 : g.nhalf - 1;
    if (threadIdx.x < kWave) {
      const double s = wave_row_sum(n, [&](int i) { return yflux_any<VL>(g, lim, p.cfly[h * P + i], p.v[h * P + i], h, i, q); });
      if (threadIdx.x == 0) bcast = s;
    }
    __syncthreads();
    const double pole = bcast * M(GMC_LE_LAT, h) / n / M(GMC_AREA_CELL, r);
    for (int i = threadIdx.x; i < n; i += NF) {
      const int64_t k = r * P + i;
      const double qc = q[k];
      p.qx[toff + k] = qc;
      p.qy[toff + k] = r == 0 ? qc - 0.5 * (pole - p.divy[k] * qc) * dt : qc + 0.5 * (pole - p.divy[k] * qc) * dt;
    }
    return;
  }
  if (it.r < 0) return;   // padding item (uniform over the block)
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), l = threadIdx.x % kWave;
  // rows h and r are uniform over the wavefront: row pointers are scalar and
  // the lanes index them with unsigned column offsets
  const int iw = wrap(it.c0 - 1 + l, n);   // column of the lane
  const unsigned uiw = iw;
  const int h = it.r - 1 + w, r = it.r + w;
  const bool row = w < kRowsPerItem && r <= g.nlat - 2;   // uniform over the wavefront
  const int i = it.c0 - 1 + l;   // cell of the lane
  const bool upd = row && l >= 1 && i < n;
  const unsigned ui = i;
  const int64_t ro = r * P;
  // the update's own loads, issued ahead of the fluxes
  double qc = 0.0, dvx = 0.0, dvy = 0.0;
  if (upd) {
    qc = ldu(q + ro, ui);
    dvx = ldu(p.divx + ro, ui);
    dvy = ldu(p.divy + ro, ui);
  }
  if (h <= g.nhalf - 1)
    sfy[w][l] = yflux_any<VL>(g, lim, ldu(p.cfly + h * P, uiw), ldu(p.v + h * P, uiw), h, iw, q);
  const double fx = row ? xflux<VL>(g, lim, ldu(p.cflx + ro, uiw), ldu(p.u + ro, uiw), iw, q + ro) : 0.0;
  const double fxw = __shfl_up(fx, 1, kWave);
  __syncthreads();
  if (upd) {
    const double lel = M(GMC_LE_LON, r), ac = M(GMC_AREA_CELL, r);
    const double ls = M(GMC_LE_LAT, r - 1), ln = M(GMC_LE_LAT, r);
    const double fs = sfy[w][l], fn = sfy[w + 1][l];
    stu(p.qx + toff + ro, ui, qc - 0.5 * ((fx - fxw) * lel / ac - dvx * qc) * dt);
    stu(p.qy + toff + ro, ui, qc - 0.5 * ((fn * ln - fs * ls) / ac - dvy * qc) * dt);
  }
}

struct OuterPtrs {
  const double *qx, *qy;   // tracer 0
  const double *q;         // level old, tracer 0
  double *qp;              // tracer 0: the updated tracer before the fixer
  const double *mfx, *mfy, *cflx, *cfly, *old_m, *m;
  int64_t tstride;
};

// outer fluxes (ffsl_mod.F90:410) and adv_run's flux-form update and /m
// (adv_mod.F90:183-235) of every cell, into qp
template <bool VL, int LIM>
__global__ void __launch_bounds__(NF, 8) k_adv_flux(AdvGrid g, Mesh M, int ntr, OuterPtrs p) {
  constexpr int lim = LIM;
  __shared__ double sfy[NF / kWave][kWave];
  __shared__ double bcast;
  const FluxItem it = flux_item(g, ntr);
  const int n = g.nlon;
  const int64_t toff = it.t * p.tstride, P = g.pitch;
  const double *qx = p.qx + toff, *qy = p.qy + toff, *q = p.q + toff;
  double *qp = p.qp + toff;
  const double dt = g.dt;
  if (it.pole) {
    const int r = it.r, h = r == 0 ? 0 : g.nhalf - 1;
    if (threadIdx.x < kWave) {
      const double s = wave_row_sum(n, [&](int i) { return yflux_any<VL>(g, lim, p.cfly[h * P + i], p.mfy[h * P + i], h, i, qx); });
      if (threadIdx.x == 0) bcast = s;
    }
    __syncthreads();
    const double pole = bcast * M(GMC_LE_LAT, h) / n / M(GMC_AREA_CELL, r) * dt;
    for (int i = threadIdx.x; i < n; i += NF) {
      const int64_t k = r * P + i;
      const double qv = r == 0 ? p.old_m[k] * q[k] - pole : p.old_m[k] * q[k] + pole;
      qp[k] = qv / p.m[k];
    }
    return;
  }
  if (it.r < 0) return;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), l = threadIdx.x % kWave;
  const int iw = wrap(it.c0 - 1 + l, n);
  const unsigned uiw = iw;
  const int h = it.r - 1 + w, r = it.r + w;
  const bool row = w < kRowsPerItem && r <= g.nlat - 2;
  const int i = it.c0 - 1 + l;
  const bool upd = row && l >= 1 && i < n;
  const unsigned ui = i;
  const int64_t ro = r * P;
  // the update's own loads, issued ahead of the fluxes
  double om = 0.0, qo = 0.0, mn = 1.0;
  if (upd) {
    om = ldu(p.old_m + ro, ui);
    qo = ldu(q + ro, ui);
    mn = ldu(p.m + ro, ui);
  }
  if (h <= g.nhalf - 1)
    sfy[w][l] = yflux_any<VL>(g, lim, ldu(p.cfly + h * P, uiw), ldu(p.mfy + h * P, uiw), h, iw, qx);
  const double fx = row ? xflux<VL>(g, lim, ldu(p.cflx + ro, uiw), ldu(p.mfx + ro, uiw), iw, qy + ro) : 0.0;
  const double fxw = __shfl_up(fx, 1, kWave);
  __syncthreads();
  if (upd) {
    const double lel = M(GMC_LE_LON, r), ac = M(GMC_AREA_CELL, r);
    const double ls = M(GMC_LE_LAT, r - 1), ln = M(GMC_LE_LAT, r);
    const double fs = sfy[w][l], fn = sfy[w + 1][l];
    const double qv = om * qo - ((fx - fxw) * lel + (fn * ln - fs * ls)) / ac * dt;
    stu(qp + ro, ui, qv / mn);
  }
}

// the instances of the flux kernels for a run's flux type and limiter
template <bool VL>
void (*inner_kernel_lim(int lim))(AdvGrid, Mesh, int, InnerPtrs) {
  return lim == GMC_LIMITER_PD ? k_adv_inner<VL, GMC_LIMITER_PD>
       : lim == GMC_LIMITER_NONE ? k_adv_inner<VL, GMC_LIMITER_NONE> : k_adv_inner<VL, GMC_LIMITER_MONO>;
}
void (*inner_kernel(bool vl, int lim))(AdvGrid, Mesh, int, InnerPtrs) {
  return vl ? inner_kernel_lim<true>(lim) : inner_kernel_lim<false>(lim);
}
template <bool VL>
void (*outer_kernel_lim(int lim))(AdvGrid, Mesh, int, OuterPtrs) {
  return lim == GMC_LIMITER_PD ? k_adv_flux<VL, GMC_LIMITER_PD>
       : lim == GMC_LIMITER_NONE ? k_adv_flux<VL, GMC_LIMITER_NONE> : k_adv_flux<VL, GMC_LIMITER_MONO>;
}
void (*outer_kernel(bool vl, int lim))(AdvGrid, Mesh, int, OuterPtrs) {
  return vl ? outer_kernel_lim<true>(lim) : outer_kernel_lim<false>(lim);
}

struct FixPtrs {
  const double *qp;        // tracer 0
  double *qn;              // level new, tracer 0
  const double *m, *wgt;
  int64_t tstride;
  int *flags;
  unsigned long long *fixes;
  int damp_passes;
};

// positivity fixer and longitude damping of one row of one tracer
// (adv_mod.F90:236-301), block (r, t) = (blockIdx.x, blockIdx.y)
__global__ void __launch_bounds__(NA) k_adv_fix(AdvGrid g, FixPtrs p) {
  extern __shared__ double lds[];
  const int n = g.nlon, r = blockIdx.x;
  const int64_t toff = blockIdx.y * p.tstride;
  double *qn = p.qn + toff;
  double *sq = lds, *neg = lds + n, *sd = lds + 2 * n, *sb = lds + 3 * n;
  // the longitude halo of the new level still holds the values of its last
  // fill_halo: the row's own last and first values, read before overwriting
  if (threadIdx.x == 0) {
    sb[1] = qn[r * g.pitch + n - 1];
    sb[2] = qn[r * g.pitch];
  }
  for (int i = threadIdx.x; i < n; i += NA) sq[i] = p.qp[toff + r * g.pitch + i];
  __syncthreads();
  // positivity fixer (adv_mod.F90:236-258).  The reference walks the row
  // once from west to east and fixes each value that is still negative when
  // reached, from its two neighbours.  A fix never makes a neighbour
  // negative and touches only columns i-1 .. i+1, so negative columns more
  // than two apart never see each other's fixes: the originally negative
  // columns split into clusters with gaps of at most one column, and one
  // lane walks each cluster west to east, in parallel with the others.  The
  // first and last columns use the seam values (the new level's longitude
  // halo) in place of each other.
  const double *mrow = p.m + r * g.pitch;
  for (int i = threadIdx.x; i < n; i += NA) neg[i] = sq[i] < 0 ? 1.0 : 0.0;
  __syncthreads();
  {
    unsigned long long nfix = 0;
    int bad = 0;
    double west = sb[1], east = sb[2];
    for (int i0 = threadIdx.x; i0 < n; i0 += NA) {
      if (neg[i0] == 0.0 || (i0 >= 1 && neg[i0 - 1] != 0.0) || (i0 >= 2 && neg[i0 - 2] != 0.0)) continue;
      for (int i = i0;;) {
        if (sq[i] < 0) {
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
        if (i + 1 < n && neg[i + 1] != 0.0) i = i + 1;
        else if (i + 2 < n && neg[i + 2] != 0.0) i = i + 2;
        else break;
      }
    }
    if (nfix) atomicAdd(p.fixes, nfix);
    if (bad) atomicOr(p.flags, F_NEG);
  }
  __syncthreads();
  // lon_damp_on_cell, order 2 (lon_damp_mod.F90:158-241): Jacobi sweeps of
  // the row with periodic neighbours
  const double w = (r > 0 && r < g.nlat - 1) ? p.wgt[r] : 0.0;
  if (w != 0.0) {
    double *a = sq, *b = sd;
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
  double *q = nullptr, *qx = nullptr, *qy = nullptr, *qp = nullptr;  // q: [2][ntracers][nlat][pitch]
  double *d_rows = nullptr;
  int *d_flags = nullptr;
  unsigned long long *d_fixes = nullptr;
  int *h_flags = nullptr;  // pinned
  hipStream_t stream = nullptr;
  gmc::HostStage stage;
  int uv_step = 0, mf_step = 0, old = 1, new_ = 2;
  bool have_basis = false;
  bool old_m_current = false;  // old_m already equals m (copy_old_m can be skipped)
  size_t lds_bytes = 0;
  hipEvent_t *prof_ev = nullptr;
  int prof_n = 0, prof_cap = 0;
  std::vector<int> prof_kind;
  // tangent-linear / adjoint storage: work fields (5 x ntracers fields and
  // the pole sums), and the slab copies (checkpoints, then tape slots)
  double *tl_f = nullptr;
  char *tl_ws = nullptr;
  size_t tl_ws_bytes = 0;
  // the adjoint's stream (default priority) and the events that order it
  // against the segment recomputes on the object's stream
  hipStream_t astream = nullptr;
  hipEvent_t ev_fwd = nullptr, ev_join = nullptr, ev_ready[2] = {}, ev_done[2] = {};
  // coupling to a dynamics context: its stream's work before the wind copy,
  // and the copy's end for the dynamics stream to wait on
  hipEvent_t ev_dyn_in = nullptr, ev_dyn_out = nullptr;
  // set when a failed adjoint window could not put the batch back at the
  // window's end: every later call refuses with GMC_ESTATE
  bool lost = false;
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

// A launch of one of the step's kernel classes.  Under gmc_adv_kernel_times
// the next two profiling events take the kernel's own start and end stamps
// (hipExtLaunchKernelGGL fills them from the dispatch); every argument is
// converted to the kernel's parameter type first, since the extended launch
// passes the arguments' bytes as given.
template <class... P, class... A>
void adv_launch(gmc_adv *a, int kind, void (*k)(P...), dim3 grid, dim3 block, uint32_t lds, A &&...args) {
  static_assert(sizeof...(P) == sizeof...(A), "one argument per kernel parameter");
  if (a->prof_ev && a->prof_n + 2 <= a->prof_cap) {
    hipExtLaunchKernelGGL(k, grid, block, lds, a->stream, a->prof_ev[a->prof_n], a->prof_ev[a->prof_n + 1], 0,
                          static_cast<P>(args)...);
    a->prof_n += 2;
    a->prof_kind.push_back(kind);
  } else {
    hipLaunchKernelGGL(k, grid, block, lds, a->stream, static_cast<P>(args)...);
  }
}

int copy_rows(gmc_adv *a, double *dst, const double *host, int rows, int64_t ld) {
  return ahip(a, stage_copy(a->stage, a->stream, true, dst, a->g.pitch, const_cast<double *>(host), ld, a->g.nlon,
                            rows), "upload");
}

// adv_run: accumulate, maybe advance, copy_old_m (adv_mod.F90:152-312).
// With tape != null an advancing step leaves its base state (what its
// tangent and adjoint read: the slab's fields from sm to the end of qy) in the
// tape slot, at their offsets from sm (tape_bytes): qx, qy and qp are written
// there, the rest is copied after the flux kernel; *taped_old is the old
// level, 0 when the step does not advance.
// before_fix(old level), when given, is called after an advancing step's
// flux kernel and before its fixer, while the slab holds the step's base
// state (the new level's seams included).
int run_once(gmc_adv *a, int64_t time_step, bool basis, double scale, char *tape = nullptr,
             int *taped_old = nullptr, const std::function<void(int)> &before_fix = nullptr) {
  if (taped_old) *taped_old = 0;
  const AdvGrid &g = a->g;
  const Mesh M{a->d_mesh, g.nlat};
  const int ns = a->cfg.nstep;
  const bool vl = a->cfg.flux_type == GMC_FLUX_VAN_LEER;
  // the two counters advance in step (adv_batch_mod.F90:238-251, 389-401)
  const int mode = a->uv_step == 0 ? 0 : (a->uv_step == ns ? 1 : 2);
  AccumPtrs ap{a->u, a->v, a->mfx, a->mfy, a->su, a->sv, a->smfx, a->smfy,
               basis ? a->ub : nullptr, basis ? a->vb : nullptr};
  const dim3 rows((g.nlon + NA - 1) / NA, g.nlat);
  adv_launch(a, KA_ACCUM, k_adv_accum, rows, dim3(NA), 0, g, ap, mode, ns, scale);
  a->uv_step = mode == 0 ? 1 : (mode == 1 ? 0 : a->uv_step + 1);
  a->mf_step = a->uv_step;
  if (a->uv_step == 0) {
    CflPtrs cp{a->u, a->v, a->cflx, a->cfly, a->divx, a->divy, a->d_flags};
    adv_launch(a, KA_CFL, k_adv_cfl, rows, dim3(NA), 0, g, M, cp);
  }
  if (time_step > 0 && time_step % ns == 0) {
    const int64_t ts = (int64_t)g.nlat * g.pitch;
    const int ntr = a->cfg.ntracers;
    const dim3 grid(flux_blocks(g, ntr));
    // when taping, qx, qy and qp go straight to the tape slot's copies
    auto T = [&](double *p) { return tape ? (double *)(tape + ((char *)p - (char *)a->sm)) : p; };
    double *qx = T(a->qx), *qy = T(a->qy), *qp = T(a->qp);
    InnerPtrs ip{tracer_ptr(a, a->old, 0), qx, qy, a->u, a->v, a->cflx, a->cfly, a->divx, a->divy, ts};
    adv_launch(a, KA_INNER, inner_kernel(vl, a->cfg.limiter), grid, dim3(NF), 0, g, M, ntr, ip);
    OuterPtrs op{qx, qy, tracer_ptr(a, a->old, 0), qp, a->mfx, a->mfy, a->cflx, a->cfly, a->old_m, a->sm, ts};
    adv_launch(a, KA_FLUX, outer_kernel(vl, a->cfg.limiter), grid, dim3(NF), 0, g, M, ntr, op);
    if (tape) {
      // the rest of the step's base state: m, the winds, CFL numbers, split
      // divergences and old_m, and q of both levels (the new level's seams)
      const size_t len = (char *)(a->q + 2 * ts * ntr) - (char *)a->sm;
      const int rc = ahip(a, hipMemcpyAsync(tape, a->sm, len, hipMemcpyDeviceToDevice, a->stream), "tape");
      if (rc) return rc;
      if (taped_old) *taped_old = a->old;
    }
    if (before_fix) before_fix(a->old);
    FixPtrs fp{qp, tracer_ptr(a, a->new_, 0), a->sm, a->d_wgt, ts, a->d_flags, a->d_fixes, a->cfg.lon_damp_passes};
    adv_launch(a, KA_FIX, k_adv_fix, dim3(g.nlat, ntr), dim3(NA), a->lds_bytes, g, fp);
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

// ---------------------------------------------------------------------------
// Coupling to the shallow-water dynamics (time_integrate, gmcore_mod.F90:588):
// block%state(itime) of the advection is a state level of a dynamics context.

// m = (gz - gzs) / g of a dynamics state (calc_m, operators_mod.F90:527-533):
// the mass the backward pass's calc_m leaves in state(new)
__global__ void __launch_bounds__(NA) k_adv_dyn_m(AdvGrid g, const double *__restrict__ gz,
                                                 const double *__restrict__ gzs, double grav, double rgrav,
                                                 double *m) {
  const int r = blockIdx.y, i = blockIdx.x * NA + threadIdx.x;
  if (i >= g.nlon) return;
  const int64_t k = r * g.pitch + i;
  m[k] = div_rn(gz[k] - gzs[k], grav, rgrav);
}

// f = value on rows 0 .. gridDim.y - 1 of a field
__global__ void __launch_bounds__(NA) k_adv_fill(AdvGrid g, double *f, double value) {
  const int r = blockIdx.y, i = blockIdx.x * NA + threadIdx.x;
  if (i < g.nlon) f[r * g.pitch + i] = value;
}

// block%state(level) of `dyn` as the advection's state: u_f, v_f, mfx_lon,
// mfy_lat (zero when zero_mf: the mass fluxes before any operator has run)
// and m, copied on the batch's stream after the dynamics stream's work so
// far; the dynamics stream then waits for the copies.
int dyn_state(gmc_adv *a, gmc_ctx *c, int level, bool zero_mf) {
  const AdvGrid &g = a->g;
  if (c->grid.nlon != g.nlon || c->grid.nlat != g.nlat || c->grid.pitch != g.pitch)
    return aerr(a, GMC_EINVAL, "the dynamics context is on another grid");
  if (level < 1 || level > 3) return aerr(a, GMC_EINVAL, "level must be 1..3");
  int rc = GMC_OK;
  for (hipEvent_t *e : {&a->ev_dyn_in, &a->ev_dyn_out})
    if (!rc && !*e) rc = ahip(a, hipEventCreateWithFlags(e, hipEventDisableTiming), "event");
  if (!rc) rc = ahip(a, hipEventRecord(a->ev_dyn_in, c->stream), "event record");
  if (!rc) rc = ahip(a, hipStreamWaitEvent(a->stream, a->ev_dyn_in, 0), "stream wait");
  if (rc) return rc;
  double **S = c->state[level];
  const size_t full = (size_t)g.nlat * g.pitch * sizeof(double), half = (size_t)g.nhalf * g.pitch * sizeof(double);
  rc = ahip(a, hipMemcpyAsync(a->su, S[S_U_F], full, hipMemcpyDeviceToDevice, a->stream), "copy u_f");
  if (!rc) rc = ahip(a, hipMemcpyAsync(a->sv, S[S_V_F], half, hipMemcpyDeviceToDevice, a->stream), "copy v_f");
  if (!rc && zero_mf) {
    // zeros through a fill kernel rather than a memset node, so that the
    // call may be captured in a graph
    hipLaunchKernelGGL(k_adv_fill, dim3((g.nlon + NA - 1) / NA, g.nlat), dim3(NA), 0, a->stream, g, a->smfx, 0.0);
    hipLaunchKernelGGL(k_adv_fill, dim3((g.nlon + NA - 1) / NA, g.nhalf), dim3(NA), 0, a->stream, g, a->smfy, 0.0);
  } else if (!rc) {
    rc = ahip(a, hipMemcpyAsync(a->smfx, S[S_MFX_LON], full, hipMemcpyDeviceToDevice, a->stream), "copy mfx_lon");
    if (!rc) rc = ahip(a, hipMemcpyAsync(a->smfy, S[S_MFY_LAT], half, hipMemcpyDeviceToDevice, a->stream),
                       "copy mfy_lat");
  }
  if (rc) return rc;
  hipLaunchKernelGGL(k_adv_dyn_m, dim3((g.nlon + NA - 1) / NA, g.nlat), dim3(NA), 0, a->stream, g, S[S_GZ], c->gzs,
                     c->cfg.g, 1.0 / c->cfg.g, a->sm);
  a->old_m_current = false;
  rc = ahip(a, hipGetLastError(), "dynamics state");
  if (!rc) rc = ahip(a, hipEventRecord(a->ev_dyn_out, a->stream), "event record");
  if (!rc) rc = ahip(a, hipStreamWaitEvent(c->stream, a->ev_dyn_out, 0), "stream wait");
  return rc;
}

extern "C" {

int gmc_adv_create(const gmc_mesh *mesh, const gmc_adv_config *cfg, gmc_adv **out) {
  if (!out) return GMC_EINVAL;
  *out = nullptr;
  if (!mesh || !cfg || !mesh->metrics || !cfg->lon_damp_wgt) return GMC_EINVAL;
  const int nlon = mesh->nlon, nlat = mesh->nlat;
  if (nlon < 8 || nlat < 5 || (nlon & 1) || nlon > 4096 || cfg->ntracers < 1 || cfg->nstep < 1 ||
      cfg->lon_halo_width < 1 || cfg->lon_halo_width > nlon / 2 - 4 || cfg->limiter < 0 ||
      cfg->limiter > 2 || cfg->lon_damp_passes < 0 || !(cfg->dt_adv > 0) || !(cfg->radius > 0) ||
      (cfg->flux_type != GMC_FLUX_PPM && cfg->flux_type != GMC_FLUX_VAN_LEER))
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
  a->lds_bytes = sizeof(double) * (3 * (size_t)nlon + 8);
  const size_t fld = (size_t)nlat * g.pitch;
  const int ntr = cfg->ntracers;
  // 16 single fields + q (2 levels) + qx, qy, qp per tracer
  const size_t nd = fld * (16 + 5 * (size_t)ntr) + GMC_MESH_NFIELDS * (size_t)nlat + nlat +
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
  a->qp = p; p += fld * ntr;
  a->qy = p; p += fld * ntr;
  a->d_mesh = p; p += GMC_MESH_NFIELDS * (size_t)nlat;
  a->d_wgt = p; p += nlat;
  a->d_rows = p; p += (size_t)ntr * nlat;
  a->d_flags = (int *)p;
  a->d_fixes = (unsigned long long *)(p + 2);
  rc = ahip(a, hipMemcpy(a->d_mesh, mesh->metrics, sizeof(double) * GMC_MESH_NFIELDS * nlat, hipMemcpyHostToDevice), "mesh");
  if (!rc) rc = ahip(a, hipMemcpy(a->d_wgt, cfg->lon_damp_wgt, sizeof(double) * nlat, hipMemcpyHostToDevice), "weights");
  if (!rc) rc = ahip(a, hipFuncSetAttribute((const void *)k_adv_fix, hipFuncAttributeMaxDynamicSharedMemorySize, (int)a->lds_bytes), "lds");
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
  if (a->astream) (void)hipStreamSynchronize(a->astream);
  if (a->slab) (void)hipFree(a->slab);
  if (a->tl_f) (void)hipFree(a->tl_f);
  if (a->tl_ws) (void)hipFree(a->tl_ws);
  for (hipEvent_t e : {a->ev_fwd, a->ev_join, a->ev_ready[0], a->ev_ready[1], a->ev_done[0], a->ev_done[1],
                       a->ev_dyn_in, a->ev_dyn_out})
    if (e) (void)hipEventDestroy(e);
  if (a->astream) (void)hipStreamDestroy(a->astream);
  if (a->h_flags) (void)hipHostFree(a->h_flags);
  if (a->stream) (void)hipStreamDestroy(a->stream);
  stage_free(a->stage);
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
  return ahip(a, stage_copy(a->stage, a->stream, false, tracer_ptr(a, level, tracer), a->g.pitch, host, ld,
                            a->g.nlon, a->g.nlat), "download");
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
  if (a->lost) return aerr(a, GMC_ESTATE, "a failed adjoint window left the batch mid-window; create it anew");
  if (time_step < 0) return aerr(a, GMC_EINVAL, "time_step < 0");
  int rc = run_once(a, time_step, false, 1.0);
  if (!rc) rc = check_flags(a);
  return rc;
}

int gmc_adv_init_dyn(gmc_adv *a, gmc_ctx *dyn, int32_t level) {
  if (!a) return GMC_EINVAL;
  if (a->lost) return aerr(a, GMC_ESTATE, "a failed adjoint window left the batch mid-window; create it anew");
  if (!dyn || !dyn->prepared) return aerr(a, GMC_ESTATE, "the dynamics context must be prepared first");
  int rc = dyn_state(a, dyn, level, true);
  if (!rc) rc = run_once(a, 0, false, 1.0);
  if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
  return rc;
}

int gmc_adv_run_dyn(gmc_adv *a, gmc_ctx *dyn, int32_t level, int64_t time_step) {
  if (!a) return GMC_EINVAL;
  if (a->lost) return aerr(a, GMC_ESTATE, "a failed adjoint window left the batch mid-window; create it anew");
  if (!dyn) return aerr(a, GMC_EINVAL, "null dynamics context");
  if (time_step < 0) return aerr(a, GMC_EINVAL, "time_step < 0");
  if (level < 1 || level > 3 || !dyn->derived_kept[level])
    return aerr(a, GMC_ESTATE, "the level does not hold the mass fluxes of its last forward pass "
                               "(step the dynamics with gmc_keep_derived on)");
  int rc = dyn_state(a, dyn, level, false);
  if (!rc) rc = run_once(a, time_step, false, 1.0);
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
  if (a->lost) return aerr(a, GMC_ESTATE, "a failed adjoint window left the batch mid-window; create it anew");
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
  std::vector<hipEvent_t> ev((size_t)nsteps * 5 * 2);   // a start and an end stamp per launch
  int rc = GMC_OK;
  for (auto &e : ev)
    if (!rc) rc = ahip(a, hipEventCreate(&e), "event");
  if (rc) return rc;
  a->prof_ev = ev.data();
  a->prof_cap = (int)ev.size();
  a->prof_n = 0;
  a->prof_kind.clear();
  for (int s = 0; s < nsteps && !rc; s++) rc = run_once(a, time_step0 + s, true, scale[s]);
  a->prof_ev = nullptr;
  if (!rc) rc = check_flags(a);
  double acc[KA_NCLASS] = {};
  int cnt[KA_NCLASS] = {};
  for (int k = 0; !rc && 2 * k + 1 < a->prof_n; k++) {
    float ms = 0.f;
    rc = ahip(a, hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]), "elapsed");
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

// ===========================================================================
// Tangent linear and adjoint of the FFSL step with respect to the tracers,
// the winds (which do not depend on q) held.  With the winds fixed the step is
// piecewise linear in q: the slope limiters, the PPM edge clipping and the
// positivity fixer choose among linear expressions by comparisons of q.  The
// tangent is the linear map of the piece the base state selects (ties go to
// the first candidate, as in slope_jac and ppm_tan), the adjoint its
// transpose.  Each kernel reads the base state of its step from a slab
// image: the object's own slab for the tangent, which runs inside the
// forward step between its flux kernel and its fixer, or a tape slot for
// the adjoint (the winds, CFL numbers and split divergences, q of both
// levels, the new level still holding the seam values the fixer reads, qx,
// qy and the pre-fixer qp, at their slab offsets).
//
// Tangent step: t_adv_inner (tangents of qx, qy), t_adv_outer (tangent of
// qp), both on the forward's work items, t_adv_fix (the fixer on the base's
// branches, then the damping).
// Adjoint step: a_adv_fix (damping, the fixer transposed, / m), k_adv_psum2
// (pole sums), a_adv_outer (outer fluxes transposed, added into the adjoints
// of qy and qx with atomics), a_adv_inner (inner fluxes transposed, added
// into the adjoint of q).  The adjoint's sums therefore run
// in no fixed order: it is reproducible to rounding, not bit for bit.
namespace gmc {
namespace {

// d slope / d (fm1, f, fp1) of the branch the base values select (slope_*,
// ffsl_mod.F90:1175-1205); mono: the smallest of |df|, df_min, df_max, the
// first on a tie, with the minimum and maximum of (fm1, f, fp1) taken at
// their first occurrence.  Returns the slope itself, bit for bit slope()'s
// value: a tie changes at most the sign of a zero, which fsign discards.
__device__ __forceinline__ double slope_jac(int lim, double fm1, double f, double fp1, double &g0, double &g1,
                                            double &g2) {
  const double df = (fp1 - fm1) * 0.5;
  g0 = -0.5; g1 = 0.0; g2 = 0.5;
  if (lim == GMC_LIMITER_NONE) return df;
  const double a = fabs(df), s = copysign(1.0, df);
  if (lim == GMC_LIMITER_PD) {
    if (2 * f < a) { g0 = 0.0; g1 = 2 * s * (2 * f >= 0 ? 1.0 : -1.0); g2 = 0.0; return fsign(2 * f, df); }
    return df;
  }
  int lo = 0, hi = 0;
  double vlo = fm1, vhi = fm1;
  if (f < vlo) { lo = 1; vlo = f; }
  if (f > vhi) { hi = 1; vhi = f; }
  if (fp1 < vlo) { lo = 2; vlo = fp1; }
  if (fp1 > vhi) { hi = 2; vhi = fp1; }
  const double dmn = 2 * (f - vlo), dmx = 2 * (vhi - f);
  int pick = 0;
  double cur = a;
  if (dmn < cur) { pick = 1; cur = dmn; }
  if (dmx < cur) { pick = 2; cur = dmx; }
  if (pick == 0) return df;  // fsign(|df|, df) = df
  const double t = 2 * s;
  if (pick == 1) {        // s * 2 (f - min)
    g0 = lo == 0 ? -t : 0.0; g1 = lo == 1 ? 0.0 : t; g2 = lo == 2 ? -t : 0.0;
  } else {                // s * 2 (max - f)
    g0 = hi == 0 ? t : 0.0; g1 = hi == 1 ? 0.0 : -t; g2 = hi == 2 ? t : 0.0;
  }
  return fsign(cur, df);
}

// The PPM Jacobian J = d (fl, df, f6) / d f[0..4] (ppm, ffsl_mod.F90:1147-1173)
// at the base values f is applied, never formed: forward (ppm_tan, J d) in
// the tangent, transposed (ppm_adj, J^T a) in the adjoint.  Both take the
// branch the base values select; the edge clipping keeps f -+ df when |df|
// is not above |fl - f| (|fr - f|).
// J d: the tangents (dl, dd, d6) of (fl, df, f6) for the stencil tangent
// d, by the chain rule through the slopes and the edge values
__device__ __forceinline__ void ppm_tan(int lim, const double f[5], const double d[5], double &dl, double &dd,
                                        double &d6) {
  double gl[3], gc[3], gr[3];
  const double dfl = slope_jac(lim, f[0], f[1], f[2], gl[0], gl[1], gl[2]);
  const double df = slope_jac(lim, f[1], f[2], f[3], gc[0], gc[1], gc[2]);
  const double dfr = slope_jac(lim, f[2], f[3], f[4], gr[0], gr[1], gr[2]);
  const double fl0 = 0.5 * (f[1] + f[2]) + div_rn(dfl - df, 6.0, 1.0 / 6.0);
  const double fr0 = 0.5 * (f[3] + f[2]) + div_rn(df - dfr, 6.0, 1.0 / 6.0);
  const double a = fabs(df), s = copysign(1.0, df);
  const bool cl = fabs(fl0 - f[2]) < a, cr = fabs(fr0 - f[2]) < a;
  const double tl = s * (fl0 - f[2] >= 0 ? 1.0 : -1.0), tr = s * (fr0 - f[2] >= 0 ? 1.0 : -1.0);
  const double tdl = gl[0] * d[0] + gl[1] * d[1] + gl[2] * d[2];
  const double tdc = gc[0] * d[1] + gc[1] * d[2] + gc[2] * d[3];
  const double tdr = gr[0] * d[2] + gr[1] * d[3] + gr[2] * d[4];
  const double tl0 = 0.5 * (d[1] + d[2]) + (tdl - tdc) / 6.0;
  const double tr0 = 0.5 * (d[3] + d[2]) + (tdc - tdr) / 6.0;
  const double tfl = cl ? d[2] - tl * (tl0 - d[2]) : d[2] - tdc;
  const double tfr = cr ? d[2] + tr * (tr0 - d[2]) : d[2] + tdc;
  dl = tfl;
  dd = tfr - tfl;
  d6 = 6 * d[2] - 3 * (tfl + tfr);
}

// J^T (a1, a2, a3): the five stencil cells' shares of the adjoints a1, a2,
// a3 of (fl, df, f6); written through the edge values, a1 fl + a2 (fr - fl)
// + a3 (6 f - 3 (fl + fr)) = (a1 - a2 - 3 a3) fl + (a2 - 3 a3) fr + 6 a3 f
__device__ __forceinline__ void ppm_adj(int lim, const double f[5], double a1, double a2, double a3, double out[5]) {
  double gl[3], gc[3], gr[3];
  const double dfl = slope_jac(lim, f[0], f[1], f[2], gl[0], gl[1], gl[2]);
  const double df = slope_jac(lim, f[1], f[2], f[3], gc[0], gc[1], gc[2]);
  const double dfr = slope_jac(lim, f[2], f[3], f[4], gr[0], gr[1], gr[2]);
  const double fl0 = 0.5 * (f[1] + f[2]) + div_rn(dfl - df, 6.0, 1.0 / 6.0);
  const double fr0 = 0.5 * (f[3] + f[2]) + div_rn(df - dfr, 6.0, 1.0 / 6.0);
  const double a = fabs(df), s = copysign(1.0, df);
  const bool cl = fabs(fl0 - f[2]) < a, cr = fabs(fr0 - f[2]) < a;
  const double tl = s * (fl0 - f[2] >= 0 ? 1.0 : -1.0), tr = s * (fr0 - f[2] >= 0 ? 1.0 : -1.0);
  const double A = a1 - a2 - 3 * a3, B = a2 - 3 * a3;
  // fl = f - df, or f - tl (fl0 - f); fr = f + df, or f + tr (fr0 - f):
  // the adjoints of fl0, fr0, of the three slopes, and f's direct share
  const double al0 = cl ? -tl * A : 0.0, ar0 = cr ? tr * B : 0.0;
  const double adf = (cl ? 0.0 : -A) + (cr ? 0.0 : B) - al0 / 6.0 + ar0 / 6.0;
  const double adl = al0 / 6.0, adr = -ar0 / 6.0;
  const double a_f = A + B + 6 * a3 - al0 - ar0;
  out[0] = adl * gl[0];
  out[1] = adl * gl[1] + adf * gc[0] + 0.5 * al0;
  out[2] = a_f + 0.5 * (al0 + ar0) + adl * gl[2] + adf * gc[1] + adr * gr[0];
  out[3] = adf * gc[2] + adr * gr[1] + 0.5 * ar0;
  out[4] = adr * gr[2];
}

// quadrature weights of the x-flux through U point i (hflx_ppm_cell
// :831-853): flux = wu (sum of cells k0 .. k0+n-1 + w1 fl + w2 df + w3 f6 of
// the upwind cell iu), |int(c)| clamped to the halo bound as in the forward
struct XW { int iu, k0, n; double wu, w1, w2, w3; };
__device__ __forceinline__ XW x_weights(const AdvGrid &g, double c, double u, int i) {
  int ci = (int)c;
  const double cf = c - ci;
  ci = ci > g.hw ? g.hw : (ci < -g.hw ? -g.hw : ci);
  const bool up = c > 0;
  const double s1 = up ? 1 - cf : 0, s2 = up ? 1 : -cf;
  const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
  XW w;
  w.iu = wrap(up ? i - ci : i - ci + 1, g.nlon);
  w.k0 = up ? i + 1 - ci : i + 1;
  w.n = up ? ci : -ci;
  w.wu = (up ? u : -u) / c;
  w.w1 = ds1;
  w.w2 = 0.5 * ds2;
  w.w3 = ds2 / 2.0 - ds3 / 3.0;
  return w;
}
// y-flux through V point (h, i): upwind full row ju (:855-877)
struct YW { int ju; double wv, w1, w2, w3; };
__device__ __forceinline__ YW y_weights(double c, double v, int h) {
  const bool up = c > 0;
  const double s1 = up ? 1 - c : 0, s2 = up ? 1 : -c;
  const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
  return YW{up ? h : h + 1, (up ? v : -v) / c, ds1, 0.5 * ds2, ds2 / 2.0 - ds3 / 3.0};
}
// offset of full row r (beyond a pole: mirrored, column half a circle away)
__device__ __forceinline__ int64_t yidx(const AdvGrid &g, int r, int i) {
  const bool mir = r < 0 || r > g.nlat - 1;
  const int rr = r < 0 ? -r : (r > g.nlat - 1 ? 2 * (g.nlat - 1) - r : r);
  return rr * g.pitch + (mir ? wrap(i + g.half, g.nlon) : i);
}

// tangent of the x-flux through U point i of a row: base row q, tangent d
__device__ __forceinline__ double xflux_tl(const AdvGrid &g, int lim, double c, double u, int i, const double *q,
                                           const double *d) {
  if (c == 0) return 0.0;
  const int n = g.nlon;
  const XW w = x_weights(g, c, u, i);
  double sum = 0.0;
  for (int k = 0; k < w.n; k++) sum += d[wrap(w.k0 + k, n)];
  double f[5], dd[5];
#pragma unroll
  for (int t = 0; t < 5; t++) {
    const int k = wrap(w.iu - 2 + t, n);
    f[t] = q[k];
    dd[t] = d[k];
  }
  double dl, dm, d6;
  ppm_tan(lim, f, dd, dl, dm, d6);
  return w.wu * (sum + w.w1 * dl + w.w2 * dm + w.w3 * d6);
}
// tangent of the y-flux through V point (h, i): base field q, tangent d
__device__ __forceinline__ double yflux_tl(const AdvGrid &g, int lim, double c, double v, int h, int i,
                                           const double *q, const double *d) {
  if (c == 0) return 0.0;
  const YW w = y_weights(c, v, h);
  double f[5], dd[5];
#pragma unroll
  for (int t = 0; t < 5; t++) {
    const int64_t k = yidx(g, w.ju - 2 + t, i);
    f[t] = q[k];
    dd[t] = d[k];
  }
  double dl, dm, d6;
  ppm_tan(lim, f, dd, dl, dm, d6);
  return w.wv * (w.w1 * dl + w.w2 * dm + w.w3 * d6);
}
// adjoint of xflux_tl: lam * d flux / d d added into the LDS row a (every
// cell an x-flux reads lies in its own row)
__device__ __forceinline__ void xflux_ad(const AdvGrid &g, int lim, double c, double u, int i, const double *q,
                                         double *a, double lam) {
  if (c == 0 || lam == 0) return;
  const int n = g.nlon;
  const XW w = x_weights(g, c, u, i);
  const double gw = w.wu * lam;
  for (int k = 0, j = wrap(w.k0, n); k < w.n; k++, j = j + 1 == n ? 0 : j + 1) atomicAdd(&a[j], gw);
  double f[5], o[5];
  int col[5];
#pragma unroll
  for (int t = 0; t < 5; t++) {
    col[t] = wrap(w.iu - 2 + t, n);
    f[t] = q[col[t]];
  }
  ppm_adj(lim, f, gw * w.w1, gw * w.w2, gw * w.w3, o);
#pragma unroll
  for (int t = 0; t < 5; t++) atomicAdd(&a[col[t]], o[t]);
}
__device__ __forceinline__ void yflux_ad(const AdvGrid &g, int lim, double c, double v, int h, int i,
                                         const double *q, double *a, double lam) {
  if (c == 0 || lam == 0) return;
  const YW w = y_weights(c, v, h);
  const double gw = w.wv * lam;
  double f[5], o[5];
  int k[5];  // a field has fewer than 2^31 elements
#pragma unroll
  for (int t = 0; t < 5; t++) {
    k[t] = (int)yidx(g, w.ju - 2 + t, i);
    f[t] = q[k[t]];
  }
  ppm_adj(lim, f, gw * w.w1, gw * w.w2, gw * w.w3, o);
#pragma unroll
  for (int t = 0; t < 5; t++) unsafeAtomicAdd(&a[k[t]], o[t]);
}

// threads of the row-block flux adjoints
constexpr int NAD = 512;

// base state of one taped step (tracer 0 of every per-tracer field)
struct TBase {
  const double *q, *qst, *qx, *qy, *qp;
  const double *u, *v, *mfx, *mfy, *cflx, *cfly, *divx, *divy, *old_m, *m;
};
// tangent / adjoint fields of one step (tracer 0), tracer stride ts
struct TWork {
  const double *in;   // tangent of the old level / adjoint of the new level
  double *dqx, *dqy, *dqp, *out;
  const double *ps;   // pole sums [ntracers][2]
  int64_t ts;
};

__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int w = 1; w < kWave; w <<= 1) v += __shfl_xor(v, w, kWave);
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x / kWave); k++) s += red[k];
  return s;
}

// ps[t][p] = sum over the columns of the pole row p of f
__global__ void __launch_bounds__(NA) k_adv_psum2(AdvGrid g, const double *f, int64_t ts, double *ps) {
  __shared__ double red[NA / kWave];
  const int p = blockIdx.x, t = blockIdx.y, r = p == 0 ? 0 : g.nlat - 1;
  double acc = 0.0;
  for (int i = threadIdx.x; i < g.nlon; i += NA) acc += f[t * ts + r * g.pitch + i];
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) ps[2 * t + p] = s;
}

// tangents of qx, qy (ffsl_mod.F90:350-407), on the work items of
// k_adv_inner: each wavefront forms the y-flux tangents of one V row (into
// LDS, shared by the cells either side) and the x-flux tangents of one cell
// row, the west face's from the lane below; a pole row's item sums the
// tangents of its V row's fluxes itself
__global__ void __launch_bounds__(NF) t_adv_inner(AdvGrid g, Mesh M, int lim, int ntr, TBase b, TWork w) {
  __shared__ double sfy[NF / kWave][kWave];
  __shared__ double bcast;
  const FluxItem it = flux_item(g, ntr);
  const int n = g.nlon;
  const int64_t to = it.t * w.ts, P = g.pitch;
  const double *q = b.q + to, *d = w.in + to;
  const double dt = g.dt;
  if (it.pole) {
    const int r = it.r, h = r == 0 ? 0 : g.nhalf - 1;
    if (threadIdx.x < kWave) {
      const double s = wave_row_sum(n, [&](int i) { return yflux_tl(g, lim, b.cfly[h * P + i], b.v[h * P + i], h, i, q, d); });
      if (threadIdx.x == 0) bcast = s;
    }
    __syncthreads();
    const double pole = bcast * M(GMC_LE_LAT, h) / n / M(GMC_AREA_CELL, r);
    for (int i = threadIdx.x; i < n; i += NF) {
      const int64_t k = r * P + i;
      const double dc = d[k];
      w.dqx[to + k] = dc;
      w.dqy[to + k] = r == 0 ? dc - 0.5 * (pole - b.divy[k] * dc) * dt : dc + 0.5 * (pole - b.divy[k] * dc) * dt;
    }
    return;
  }
  if (it.r < 0) return;   // padding item (uniform over the block)
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), l = threadIdx.x % kWave;
  const int iw = wrap(it.c0 - 1 + l, n);   // column of the lane
  const int h = it.r - 1 + wv, r = it.r + wv;
  if (h <= g.nhalf - 1) sfy[wv][l] = yflux_tl(g, lim, b.cfly[h * P + iw], b.v[h * P + iw], h, iw, q, d);
  const bool row = wv < kRowsPerItem && r <= g.nlat - 2;   // uniform over the wavefront
  const double fx = row ? xflux_tl(g, lim, b.cflx[r * P + iw], b.u[r * P + iw], iw, q + r * P, d + r * P) : 0.0;
  const double fxw = __shfl_up(fx, 1, kWave);
  __syncthreads();
  const int i = it.c0 - 1 + l;   // cell of the lane
  if (row && l >= 1 && i < n) {
    const double lel = M(GMC_LE_LON, r), ac = M(GMC_AREA_CELL, r);
    const double ls = M(GMC_LE_LAT, r - 1), ln = M(GMC_LE_LAT, r);
    const double fs = sfy[wv][l], fn = sfy[wv + 1][l];
    const int64_t k = r * P + i;
    const double dc = d[k];
    w.dqx[to + k] = dc - 0.5 * ((fx - fxw) * lel / ac - b.divx[k] * dc) * dt;
    w.dqy[to + k] = dc - 0.5 * ((fn * ln - fs * ls) / ac - b.divy[k] * dc) * dt;
  }
}

// tangent of the outer fluxes and the update (adv_mod.F90:183-235) into
// dqp, on the work items of k_adv_flux
__global__ void __launch_bounds__(NF) t_adv_outer(AdvGrid g, Mesh M, int lim, int ntr, TBase b, TWork w) {
  __shared__ double sfy[NF / kWave][kWave];
  __shared__ double bcast;
  const FluxItem it = flux_item(g, ntr);
  const int n = g.nlon;
  const int64_t to = it.t * w.ts, P = g.pitch;
  const double *qx = b.qx + to, *qy = b.qy + to, *dqx = w.dqx + to, *dqy = w.dqy + to, *d = w.in + to;
  const double dt = g.dt;
  if (it.pole) {
    const int r = it.r, h = r == 0 ? 0 : g.nhalf - 1;
    if (threadIdx.x < kWave) {
      const double s = wave_row_sum(n, [&](int i) { return yflux_tl(g, lim, b.cfly[h * P + i], b.mfy[h * P + i], h, i, qx, dqx); });
      if (threadIdx.x == 0) bcast = s;
    }
    __syncthreads();
    const double pole = bcast * M(GMC_LE_LAT, h) / n / M(GMC_AREA_CELL, r) * dt;
    for (int i = threadIdx.x; i < n; i += NF) {
      const int64_t k = r * P + i;
      const double num = r == 0 ? b.old_m[k] * d[k] - pole : b.old_m[k] * d[k] + pole;
      w.dqp[to + k] = num / b.m[k];
    }
    return;
  }
  if (it.r < 0) return;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), l = threadIdx.x % kWave;
  const int iw = wrap(it.c0 - 1 + l, n);
  const int h = it.r - 1 + wv, r = it.r + wv;
  if (h <= g.nhalf - 1) sfy[wv][l] = yflux_tl(g, lim, b.cfly[h * P + iw], b.mfy[h * P + iw], h, iw, qx, dqx);
  const bool row = wv < kRowsPerItem && r <= g.nlat - 2;
  const double fx = row ? xflux_tl(g, lim, b.cflx[r * P + iw], b.mfx[r * P + iw], iw, qy + r * P, dqy + r * P) : 0.0;
  const double fxw = __shfl_up(fx, 1, kWave);
  __syncthreads();
  const int i = it.c0 - 1 + l;
  if (row && l >= 1 && i < n) {
    const double lel = M(GMC_LE_LON, r), ac = M(GMC_AREA_CELL, r);
    const double ls = M(GMC_LE_LAT, r - 1), ln = M(GMC_LE_LAT, r);
    const double fs = sfy[wv][l], fn = sfy[wv + 1][l];
    const int64_t k = r * P + i;
    const double num = b.old_m[k] * d[k] - ((fx - fxw) * lel + (fn * ln - fs * ls)) / ac * dt;
    w.dqp[to + k] = num / b.m[k];
  }
}

// One row's positivity fixer (adv_mod.F90:236-258) on the base values sb,
// carrying the tangent sd (when sd != null) or recording each cell's branch
// in code (0 none, 1 both neighbours, 2 west, 3 east, 4 none possible) for
// the adjoint.  Clusters of originally negative columns (gaps of at most one
// column) are independent, one lane each, walked west to east as in
// k_adv_fix; with code_end != null, code_end[i0] receives the last column
// walked from cluster start i0.
__device__ __forceinline__ void fixer_walk(int n, double *sb, double *sd, const double *neg, double *code,
                                           double *code_end, const double *mrow, double west, double east) {
  for (int i0 = threadIdx.x; i0 < n; i0 += NA) {
    if (neg[i0] == 0.0 || (i0 >= 1 && neg[i0 - 1] != 0.0) || (i0 >= 2 && neg[i0 - 2] != 0.0)) continue;
    int i = i0;
    for (;;) {
      double c = 0.0;
      if (sb[i] < 0) {
        const int im = i == 0 ? n - 1 : i - 1, ip = i == n - 1 ? 0 : i + 1;
        const double qm0 = sb[i] * mrow[i];
        const double qm1 = (i == 0 ? west : sb[im]) * mrow[im];
        const double qm2 = (i == n - 1 ? east : sb[ip]) * mrow[ip];
        const double qm0_half = 0.5 * qm0;
        const double d0 = sd ? sd[i] * mrow[i] : 0.0;
        if (qm1 >= qm0_half && qm2 >= qm0_half) {
          c = 1.0;
          const double a = (qm1 - qm0_half) / mrow[im], bb = (qm2 - qm0_half) / mrow[ip];
          if (i == 0) west = a; else sb[im] = a;
          if (i == n - 1) east = bb; else sb[ip] = bb;
          if (sd && i > 0) sd[im] = (sd[im] * mrow[im] - 0.5 * d0) / mrow[im];
          if (sd && i < n - 1) sd[ip] = (sd[ip] * mrow[ip] - 0.5 * d0) / mrow[ip];
        } else if (qm1 > qm0) {
          c = 2.0;
          const double a = (qm1 - qm0) / mrow[im];
          if (i == 0) west = a; else sb[im] = a;
          if (sd && i > 0) sd[im] = (sd[im] * mrow[im] - d0) / mrow[im];
        } else if (qm2 > qm0) {
          c = 3.0;
          const double bb = (qm2 - qm0) / mrow[ip];
          if (i == n - 1) east = bb; else sb[ip] = bb;
          if (sd && i < n - 1) sd[ip] = (sd[ip] * mrow[ip] - d0) / mrow[ip];
        } else {
          c = 4.0;
        }
        sb[i] = 0;
        if (sd) sd[i] = 0;
      }
      if (code) code[i] = c;
      if (i + 1 < n && neg[i + 1] != 0.0) i = i + 1;
      else if (i + 2 < n && neg[i + 2] != 0.0) i = i + 2;
      else break;
    }
    if (code_end) code_end[i0] = (double)i;
  }
}

// Jacobi sweeps of lon_damp_on_cell, order 2 (lon_damp_mod.F90:158-241), on
// the LDS row a (b scratch); returns the row holding the result.  The sweep
// is a symmetric circulant, its own transpose.
__device__ __forceinline__ double *damp_row(int n, double w, int passes, double *a, double *b) {
  if (w == 0.0) return a;
  for (int pass = 0; pass < passes; pass++) {
    for (int i = threadIdx.x; i < n; i += NA)
      b[i] = w * (a[i] + 0.25 * (a[wrap(i - 1, n)] - 2.0 * a[i] + a[wrap(i + 1, n)])) + (1 - w) * a[i];
    __syncthreads();
    double *t = a; a = b; b = t;
  }
  return a;
}

// tangent of the fixer and the damping of row r of tracer t: dqp -> out
__global__ void __launch_bounds__(NA) t_adv_fix(AdvGrid g, TBase b, TWork w, const double *wgt, int passes) {
  extern __shared__ double lds[];
  const int n = g.nlon, r = blockIdx.x;
  const int64_t to = blockIdx.y * w.ts, row = to + r * g.pitch;
  double *sb = lds, *sd = lds + n, *st = lds + 2 * n, *neg = lds + 3 * n;
  const double *mrow = b.m + r * g.pitch;
  for (int i = threadIdx.x; i < n; i += NA) {
    sb[i] = b.qp[row + i];
    sd[i] = w.dqp[row + i];
    neg[i] = sb[i] < 0 ? 1.0 : 0.0;
  }
  __syncthreads();
  fixer_walk(n, sb, sd, neg, nullptr, nullptr, mrow, b.qst[row + n - 1], b.qst[row]);
  __syncthreads();
  const double wr = (r > 0 && r < g.nlat - 1) ? wgt[r] : 0.0;
  const double *res = damp_row(n, wr, passes, sd, st);
  for (int i = threadIdx.x; i < n; i += NA) w.out[row + i] = res[i];
}

// adjoint of t_adv_fix and of the division by m: the adjoint of the new
// level (in) -> dqp holds the adjoint of the update's numerator; clears the
// rows of dqx the y-flux adjoints accumulate into
__global__ void __launch_bounds__(NA) a_adv_fix(AdvGrid g, TBase b, TWork w, const double *wgt, int passes) {
  extern __shared__ double lds[];
  const int n = g.nlon, r = blockIdx.x;
  const int64_t to = blockIdx.y * w.ts, row = to + r * g.pitch;
  double *sl = lds, *st = lds + n, *sb = lds + 2 * n, *neg = lds + 3 * n, *code = lds + 4 * n;
  double *cend = st;  // the damping scratch is free again once it is done
  const double *mrow = b.m + r * g.pitch;
  for (int i = threadIdx.x; i < n; i += NA) {
    sl[i] = w.in[row + i];
    sb[i] = b.qp[row + i];
    neg[i] = sb[i] < 0 ? 1.0 : 0.0;
    code[i] = 0.0;
    w.dqx[row + i] = 0.0;
  }
  __syncthreads();
  const double wr = (r > 0 && r < g.nlat - 1) ? wgt[r] : 0.0;
  double *lam = damp_row(n, wr, passes, sl, st);
  if (lam != sl) {
    for (int i = threadIdx.x; i < n; i += NA) sl[i] = lam[i];
    __syncthreads();
  }
  fixer_walk(n, sb, nullptr, neg, code, cend, mrow, b.qst[row + n - 1], b.qst[row]);
  __syncthreads();
  // the fixer transposed, each cluster east to west: the cell's own
  // adjoint is replaced by what its neighbours' adjoints take from it; the
  // seam writes were discarded, so they contribute nothing
  for (int i0 = threadIdx.x; i0 < n; i0 += NA) {
    if (neg[i0] == 0.0 || (i0 >= 1 && neg[i0 - 1] != 0.0) || (i0 >= 2 && neg[i0 - 2] != 0.0)) continue;
    for (int i = (int)cend[i0]; i >= i0; i--) {
      const int c = (int)code[i];
      if (c == 0) continue;
      const double f = c == 1 ? 0.5 : 1.0;
      double gsum = 0.0;
      if ((c == 1 || c == 2) && i > 0) gsum -= f * mrow[i] / mrow[i - 1] * sl[i - 1];
      if ((c == 1 || c == 3) && i < n - 1) gsum -= f * mrow[i] / mrow[i + 1] * sl[i + 1];
      sl[i] = gsum;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += NA) w.dqp[row + i] = sl[i] / mrow[i];
}

// adjoint of the update and the outer fluxes, block (row r, tracer t): dqp
// (numerator adjoint) -> out = old_m * dqp (the direct term); the x-flux
// adjoints gathered in LDS and stored as row r of dqy (nothing else adds
// there), the y-flux adjoints added into dqx with atomics
__global__ void __launch_bounds__(NAD) a_adv_outer(AdvGrid g, Mesh M, int lim, TBase b, TWork w) {
  extern __shared__ double acc[];
  const int n = g.nlon, r = blockIdx.x, t = blockIdx.y;
  const int64_t to = t * w.ts, P = g.pitch;
  const double dt = g.dt;
  const double *an = w.dqp + to;
  for (int i = threadIdx.x; i < n; i += NAD) acc[i] = 0.0;
  // the row's factors: x-flux, and the V row h = r's shares of the cell
  // rows either side (a pole row through its zonal sum)
  const bool inner = r >= 1 && r <= g.nlat - 2;
  const double kx = inner ? M(GMC_LE_LON, r) / M(GMC_AREA_CELL, r) * dt : 0.0;
  const int h = r;
  const bool vrow = h <= g.nlat - 2;
  const double ks = vrow && h >= 1 ? M(GMC_LE_LAT, h) / M(GMC_AREA_CELL, h) * dt : 0.0;
  const double kn = vrow && h + 1 <= g.nlat - 2 ? M(GMC_LE_LAT, h) / M(GMC_AREA_CELL, h + 1) * dt : 0.0;
  double kp = 0.0;
  if (vrow && h == 0) kp -= M(GMC_LE_LAT, 0) / g.nlon / M(GMC_AREA_CELL, 0) * dt * w.ps[2 * t];
  if (vrow && h == g.nlat - 2) kp += M(GMC_LE_LAT, h) / g.nlon / M(GMC_AREA_CELL, g.nlat - 1) * dt * w.ps[2 * t + 1];
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += NAD) {
    const int64_t k = r * P + i;
    w.out[to + k] = b.old_m[k] * an[k];
    if (inner) xflux_ad(g, lim, b.cflx[k], b.mfx[k], i, b.qy + to + r * P, acc, (an[r * P + wrap(i + 1, n)] - an[k]) * kx);
    if (!vrow) continue;
    double lf = kp;
    if (h >= 1) lf -= ks * an[h * P + i];
    if (h + 1 <= g.nlat - 2) lf += kn * an[(h + 1) * P + i];
    yflux_ad(g, lim, b.cfly[h * P + i], b.mfy[h * P + i], h, i, b.qx + to, w.dqx + to, lf);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += NAD) w.dqy[to + r * P + i] = acc[i];
}

// adjoint of qx, qy and the inner fluxes, block (row r, tracer t): the
// direct terms and the x-flux adjoints gathered in LDS and added into row r
// of out once per cell, the y-flux adjoints added into out with atomics
// (the pole sums in w.ps are not read)
__global__ void __launch_bounds__(NAD) a_adv_inner(AdvGrid g, Mesh M, int lim, TBase b, TWork w) {
  extern __shared__ double acc[];
  const int n = g.nlon, r = blockIdx.x, t = blockIdx.y;
  const int64_t to = t * w.ts, P = g.pitch;
  const double dt = g.dt;
  const double *ax = w.dqx + to, *ay = w.dqy + to;
  for (int i = threadIdx.x; i < n; i += NAD) acc[i] = 0.0;
  const bool inner = r >= 1 && r <= g.nlat - 2;
  const double kx = inner ? 0.5 * dt * M(GMC_LE_LON, r) / M(GMC_AREA_CELL, r) : 0.0;
  const double ks = r >= 1 && r <= g.nlat - 2 ? 0.5 * dt * M(GMC_LE_LAT, r) / M(GMC_AREA_CELL, r) : 0.0;
  const double kn = r + 1 <= g.nlat - 2 ? 0.5 * dt * M(GMC_LE_LAT, r) / M(GMC_AREA_CELL, r + 1) : 0.0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += NAD) {
    const int64_t k = r * P + i;
    const double gx = ax[k], gy = ay[k];
    double add;
    if (r == 0) add = gx + (1 + 0.5 * b.divy[k] * dt) * gy;
    else if (r == g.nlat - 1) add = gx + (1 - 0.5 * b.divy[k] * dt) * gy;
    else add = (1 + 0.5 * b.divx[k] * dt) * gx + (1 + 0.5 * b.divy[k] * dt) * gy;
    atomicAdd(&acc[i], add);
    if (inner) xflux_ad(g, lim, b.cflx[k], b.u[k], i, b.q + to + r * P, acc, kx * (ax[r * P + wrap(i + 1, n)] - gx));
    if (r > g.nlat - 2) continue;
    const int h = r;
    double lf = 0.0;
    if (h >= 1) lf -= ks * ay[h * P + i];
    if (h + 1 <= g.nlat - 2) lf += kn * ay[(h + 1) * P + i];
    // the pole rows' qy feed no outer flux (the outer x-fluxes skip the
    // pole rows), so their adjoint, and the pole-sum terms it would bring
    // into the V rows next to the poles, are zero
    yflux_ad(g, lim, b.cfly[h * P + i], b.v[h * P + i], h, i, b.q + to, w.out + to, lf);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += NAD) unsafeAtomicAdd(&w.out[to + r * P + i], acc[i]);
}

// misfit of the window's end state: d = q - y in place of y (every tracer),
// rows[t][r] = sum over the row of d^2; block (r, t)
__global__ void __launch_bounds__(NA) k_adv_misfit(AdvGrid g, const double *q, double *d, int64_t ts,
                                                   double *rows) {
  __shared__ double red[NA / kWave];
  const int r = blockIdx.x, t = blockIdx.y;
  const int64_t row = t * ts + r * g.pitch;
  double acc = 0.0;
  for (int i = threadIdx.x; i < g.nlon; i += NA) {
    const double e = q[row + i] - d[row + i];
    d[row + i] = e;
    acc += e * e;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) rows[t * g.nlat + r] = s;
}

}  // namespace
}  // namespace gmc

namespace {

// host-side part of a checkpoint (the device part is a copy of the slab)
struct HostCk { int uv_step, mf_step, old, new_; bool old_m_current; };
HostCk host_ck(const gmc_adv *a) { return HostCk{a->uv_step, a->mf_step, a->old, a->new_, a->old_m_current}; }
void host_restore(gmc_adv *a, const HostCk &c) {
  a->uv_step = c.uv_step; a->mf_step = c.mf_step; a->old = c.old; a->new_ = c.new_;
  a->old_m_current = c.old_m_current;
}

size_t slot_bytes(const gmc_adv *a) { return (a->slab_bytes + 255) / 256 * 256; }
size_t tl_fields(const gmc_adv *a) { return (size_t)a->cfg.ntracers * a->g.nlat * a->g.pitch; }
// a tape slot: the slab's fields from sm to the end of qy (m, the accumulated
// winds, CFL numbers, split divergences, old_m, q of both levels, qx, qp,
// qy), not the state winds, the mesh and the counters
size_t tape_bytes(const gmc_adv *a) {
  return ((size_t)((const char *)(a->qy + tl_fields(a)) - (const char *)a->sm) + 255) / 256 * 256;
}

// work fields (5 per tracer: two tangent / adjoint levels, dqx, dqy, dqp,
// then 4 pole sums per tracer), nck slab copies (checkpoints) and ntape tape
// slots after them
int tl_alloc(gmc_adv *a, int nck, int ntape) {
  int rc = GMC_OK;
  if (!a->tl_f) {
    const size_t bytes = (5 * tl_fields(a) + 4 * (size_t)a->cfg.ntracers) * sizeof(double);
    rc = ahip(a, hipMalloc((void **)&a->tl_f, bytes), "tangent-linear work fields");
    if (!rc) rc = ahip(a, hipMemset(a->tl_f, 0, bytes), "memset");
    const int lds = (int)(5 * (size_t)a->g.nlon * sizeof(double));
    if (!rc) rc = ahip(a, hipFuncSetAttribute((const void *)t_adv_fix, hipFuncAttributeMaxDynamicSharedMemorySize, lds), "lds");
    if (!rc) rc = ahip(a, hipFuncSetAttribute((const void *)a_adv_fix, hipFuncAttributeMaxDynamicSharedMemorySize, lds), "lds");
    if (rc) return rc;
  }
  const size_t need = (size_t)nck * slot_bytes(a) + (size_t)ntape * tape_bytes(a);
  if (need > a->tl_ws_bytes) {
    if (a->tl_ws) {
      (void)hipStreamSynchronize(a->stream);
      (void)hipFree(a->tl_ws);
      a->tl_ws = nullptr;
      a->tl_ws_bytes = 0;
    }
    rc = ahip(a, hipMalloc((void **)&a->tl_ws, need), "tangent-linear checkpoints and tapes");
    if (!rc) a->tl_ws_bytes = need;
  }
  return rc;
}
char *tl_slot(gmc_adv *a, int k) { return a->tl_ws + (size_t)k * slot_bytes(a); }
char *tape_slot(gmc_adv *a, int nck, int j) { return a->tl_ws + (size_t)nck * slot_bytes(a) + (size_t)j * tape_bytes(a); }
int slab_to(gmc_adv *a, char *dst) {
  return ahip(a, hipMemcpyAsync(dst, a->slab, a->slab_bytes, hipMemcpyDeviceToDevice, a->stream), "checkpoint");
}
int slab_from(gmc_adv *a, const char *src) {
  return ahip(a, hipMemcpyAsync(a->slab, src, a->slab_bytes, hipMemcpyDeviceToDevice, a->stream), "restore");
}
// the fields only (not the mesh, weights and counters after them, which a
// concurrent adjoint reads)
int fields_from(gmc_adv *a, const char *src) {
  const size_t len = (char *)a->d_mesh - (char *)a->slab;
  return ahip(a, hipMemcpyAsync(a->slab, src, len, hipMemcpyDeviceToDevice, a->stream), "restore");
}
int adj_stream(gmc_adv *a) {
  if (a->astream) return GMC_OK;
  int rc = ahip(a, hipStreamCreateWithFlags(&a->astream, hipStreamNonBlocking), "adjoint stream");
  for (hipEvent_t *e : {&a->ev_fwd, &a->ev_join, &a->ev_ready[0], &a->ev_ready[1], &a->ev_done[0], &a->ev_done[1]})
    if (!rc) rc = ahip(a, hipEventCreateWithFlags(e, hipEventDisableTiming), "event");
  return rc;
}

// the fields of a tape slot, or of the live slab with slot = (char *)a->sm
TBase tape_base(gmc_adv *a, const char *slot, int old) {
  auto S = [&](const double *p) { return (const double *)(slot + ((const char *)p - (const char *)a->sm)); };
  return TBase{S(tracer_ptr(a, old, 0)), S(tracer_ptr(a, 3 - old, 0)), S(a->qx), S(a->qy), S(a->qp),
               S(a->u), S(a->v), S(a->mfx), S(a->mfy), S(a->cflx), S(a->cfly), S(a->divx), S(a->divy),
               S(a->old_m), S(a->sm)};
}

// tangent of one taped step: in (old level) -> out (new level)
void tl_step(gmc_adv *a, const char *slot, int old, const double *in, double *out) {
  const AdvGrid &g = a->g;
  const Mesh M{a->d_mesh, g.nlat};
  const int ntr = a->cfg.ntracers, lim = a->cfg.limiter;
  const int64_t ts = (int64_t)g.nlat * g.pitch;
  const size_t F = tl_fields(a);
  double *dqx = a->tl_f + 2 * F, *dqy = a->tl_f + 3 * F, *dqp = a->tl_f + 4 * F, *ps = a->tl_f + 5 * F;
  const TBase b = tape_base(a, slot, old);
  const dim3 items(flux_blocks(g, ntr));
  hipLaunchKernelGGL(t_adv_inner, items, dim3(NF), 0, a->stream, g, M, lim, ntr, b, TWork{in, dqx, dqy, dqp, out, ps, ts});
  hipLaunchKernelGGL(t_adv_outer, items, dim3(NF), 0, a->stream, g, M, lim, ntr, b, TWork{in, dqx, dqy, dqp, out, ps, ts});
  hipLaunchKernelGGL(t_adv_fix, dim3(g.nlat, ntr), dim3(NA), 4 * (size_t)g.nlon * sizeof(double), a->stream, g, b,
                     TWork{in, dqx, dqy, dqp, out, ps, ts}, a->d_wgt, a->cfg.lon_damp_passes);
}

// adjoint of one taped step on stream st: in (adjoint of the new level) ->
// out (of the old level)
void ad_step(gmc_adv *a, const char *slot, int old, const double *in, double *out, hipStream_t st) {
  const AdvGrid &g = a->g;
  const Mesh M{a->d_mesh, g.nlat};
  const int ntr = a->cfg.ntracers, lim = a->cfg.limiter;
  const int64_t ts = (int64_t)g.nlat * g.pitch;
  const size_t F = tl_fields(a);
  double *dqx = a->tl_f + 2 * F, *dqy = a->tl_f + 3 * F, *dqp = a->tl_f + 4 * F, *ps = a->tl_f + 5 * F;
  const TBase b = tape_base(a, slot, old);
  const dim3 rows(g.nlat, ntr), poles(2, ntr);
  hipLaunchKernelGGL(a_adv_fix, dim3(g.nlat, ntr), dim3(NA), 5 * (size_t)g.nlon * sizeof(double), st, g, b,
                     TWork{in, dqx, dqy, dqp, out, ps, ts}, a->d_wgt, a->cfg.lon_damp_passes);
  hipLaunchKernelGGL(k_adv_psum2, poles, dim3(NA), 0, st, g, dqp, ts, ps);
  const size_t lrow = (size_t)g.nlon * sizeof(double);
  hipLaunchKernelGGL(a_adv_outer, rows, dim3(NAD), lrow, st, g, M, lim, b, TWork{in, dqx, dqy, dqp, out, ps, ts});
  hipLaunchKernelGGL(a_adv_inner, rows, dim3(NAD), lrow, st, g, M, lim, b, TWork{in, dqx, dqy, dqp, out, ps, ts});
}

int tl_check(gmc_adv *a, const double *scale, int32_t nsteps, int64_t time_step0, const double *buf, int64_t ld) {
  if (a->lost) return aerr(a, GMC_ESTATE, "a failed adjoint window left the batch mid-window; create it anew");
  if (!a->have_basis) return aerr(a, GMC_ESTATE, "gmc_adv_set_wind_basis must come first");
  if (a->cfg.flux_type != GMC_FLUX_PPM)
    return aerr(a, GMC_EINVAL, "the tangent linear and adjoint cover the PPM flux only");
  if (!scale || nsteps < 0 || time_step0 < 0 || !buf || ld < a->g.nlon)
    return aerr(a, GMC_EINVAL, "bad tangent-linear / adjoint arguments");
  return GMC_OK;
}

// [ntracers][nlat][ld] host <-> a device tangent / adjoint level
int tl_copy(gmc_adv *a, bool up, double *dev, double *host, int64_t ld) {
  const int rows = a->cfg.ntracers * a->g.nlat;
  int rc = ahip(a, stage_copy(a->stage, a->stream, up, dev, a->g.pitch, host, ld, a->g.nlon, rows),
                up ? "upload" : "download");
  if (!rc) rc = ahip(a, hipGetLastError(), "tangent-linear / adjoint launch");
  return rc;
}

}  // namespace

extern "C" {

int gmc_adv_tl_drive(gmc_adv *a, int64_t time_step0, const double *scale, int32_t nsteps, double *dq, int64_t ld) {
  if (!a) return GMC_EINVAL;
  int rc = tl_check(a, scale, nsteps, time_step0, dq, ld);
  if (!rc) rc = tl_alloc(a, 0, 0);
  if (rc) return rc;
  const size_t F = tl_fields(a);
  double *cur = a->tl_f, *nxt = a->tl_f + F;
  rc = tl_copy(a, true, cur, dq, ld);
  // each advancing step's tangent runs on the live slab, between the step's
  // flux kernel and its fixer
  const auto tangent = [&](int old) {
    tl_step(a, (const char *)a->sm, old, cur, nxt);
    std::swap(cur, nxt);
  };
  for (int s = 0; s < nsteps && !rc; s++) {
    rc = run_once(a, time_step0 + s, true, scale[s], nullptr, nullptr, tangent);
    if (!rc) rc = ahip(a, hipGetLastError(), "tangent-linear launch");
  }
  if (!rc) rc = check_flags(a);
  if (!rc) rc = tl_copy(a, false, cur, dq, ld);
  if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
  return rc;
}

}  // extern "C"

namespace {

// the adjoint of the window, forced at its end by lam (y == null) or by the
// misfit q_N - y of J = 1/2 sum (q_N - y)^2 (then *cost = J); the result
// goes to out
int adv_adjoint(gmc_adv *a, int64_t time_step0, const double *scale, int32_t nsteps, int32_t ckpt_every,
                const double *lam, const double *y, double *out, double *cost, int64_t ld) {
  if (!a) return GMC_EINVAL;
  int rc = tl_check(a, scale, nsteps, time_step0, y ? y : lam, ld);
  if (!rc && !out) rc = aerr(a, GMC_EINVAL, "null output");
  if (!rc && ckpt_every < 1) rc = aerr(a, GMC_EINVAL, "ckpt_every must be at least 1");
  if (rc) return rc;
  const int K = std::min<int>(ckpt_every, std::max<int>(nsteps, 1));
  const int nseg = (nsteps + K - 1) / K;
  // slots: checkpoints 0 .. nseg (the last is the window's end), then two
  // tape buffers (one when there is a single segment) of as many slots as a
  // segment has advancing steps (adv_run advances at time steps k nstep,
  // k > 0); segment seg is taped into buffer seg % 2
  const auto advances = [&](int s) { const int64_t t = time_step0 + s; return t > 0 && t % a->cfg.nstep == 0; };
  int amax = 0;
  for (int seg = 0; seg < nseg; seg++) {
    int c = 0;
    for (int s = seg * K; s < std::min(nsteps, (seg + 1) * K); s++) c += advances(s);
    amax = std::max(amax, c);
  }
  rc = tl_alloc(a, nseg + 1, (nseg > 1 ? 2 : 1) * amax);
  if (!rc) rc = adj_stream(a);
  if (rc) return rc;
  // the tape slot of step s of segment seg: its rank among the segment's
  // advancing steps (none for a step that does not advance)
  auto tape = [&](int seg, int s) -> char * {
    if (!advances(s)) return nullptr;
    int j = 0;
    for (int q = seg * K; q < s; q++) j += advances(q);
    return tape_slot(a, nseg + 1, (seg % 2) * amax + j);
  };
  std::vector<HostCk> hck(nseg + 1);
  std::vector<int> olds[2] = {std::vector<int>(K), std::vector<int>(K)};
  const int slast = (nseg - 1) * K;  // the last segment is taped on the way
  for (int s = 0; s < nsteps && !rc; s++) {
    if (s % K == 0) {
      hck[s / K] = host_ck(a);
      rc = slab_to(a, tl_slot(a, s / K));
    }
    if (!rc && s >= slast)
      rc = run_once(a, time_step0 + s, true, scale[s], tape(nseg - 1, s), &olds[(nseg - 1) % 2][s - slast]);
    else if (!rc)
      rc = run_once(a, time_step0 + s, true, scale[s]);
  }
  bool window_done = false;   // slot nseg holds the window's end
  if (!rc) {
    hck[nseg] = host_ck(a);
    rc = slab_to(a, tl_slot(a, nseg));
    window_done = !rc;
  }
  if (!rc) rc = check_flags(a);
  const size_t F = tl_fields(a);
  double *cur = a->tl_f, *nxt = a->tl_f + F;
  if (!rc) rc = tl_copy(a, true, cur, const_cast<double *>(y ? y : lam), ld);
  if (!rc && y) {
    const AdvGrid &g = a->g;
    const int ntr = a->cfg.ntracers;
    hipLaunchKernelGGL(k_adv_misfit, dim3(g.nlat, ntr), dim3(NA), 0, a->stream, g, tracer_ptr(a, a->old, 0), cur,
                       (int64_t)g.nlat * g.pitch, a->d_rows);
    std::vector<double> rows((size_t)ntr * g.nlat);
    rc = ahip(a, hipGetLastError(), "misfit launch");
    if (!rc) rc = ahip(a, hipMemcpyAsync(rows.data(), a->d_rows, rows.size() * sizeof(double), hipMemcpyDeviceToHost,
                                         a->stream), "misfit copy");
    if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
    double j = 0.0;
    for (double v : rows) j += v;
    if (cost) *cost = 0.5 * j;
  }
  // Last segment first: the adjoint of segment seg runs on the adjoint
  // stream from tape buffer seg % 2 while the object's stream recomputes
  // segment seg - 1 from its checkpoint into the other buffer, once the
  // adjoint of segment seg + 1 has finished reading it.
  if (!rc) rc = ahip(a, hipEventRecord(a->ev_fwd, a->stream), "event");
  if (!rc) rc = ahip(a, hipStreamWaitEvent(a->astream, a->ev_fwd, 0), "wait");
  for (int seg = nseg - 1; seg >= 0 && !rc; seg--) {
    const int b = seg % 2;
    if (seg > 0) {
      const int p0 = (seg - 1) * K, p1 = seg * K;
      if (seg + 1 <= nseg - 1) rc = ahip(a, hipStreamWaitEvent(a->stream, a->ev_done[1 - b], 0), "wait");
      host_restore(a, hck[seg - 1]);
      if (!rc) rc = fields_from(a, tl_slot(a, seg - 1));
      for (int s = p0; s < p1 && !rc; s++)
        rc = run_once(a, time_step0 + s, true, scale[s], tape(seg - 1, s), &olds[1 - b][s - p0]);
      if (!rc) rc = ahip(a, hipEventRecord(a->ev_ready[1 - b], a->stream), "event");
    }
    if (!rc && seg < nseg - 1) rc = ahip(a, hipStreamWaitEvent(a->astream, a->ev_ready[b], 0), "wait");
    const int s0 = seg * K, s1 = std::min(nsteps, s0 + K);
    for (int s = s1 - 1; s >= s0 && !rc; s--) {
      if (!olds[b][s - s0]) continue;
      ad_step(a, tape(seg, s), olds[b][s - s0], cur, nxt, a->astream);
      std::swap(cur, nxt);
      rc = ahip(a, hipGetLastError(), "adjoint launch");
    }
    if (!rc) rc = ahip(a, hipEventRecord(a->ev_done[b], a->astream), "event");
  }
  if (!rc) rc = ahip(a, hipEventRecord(a->ev_join, a->astream), "event");
  if (!rc) rc = ahip(a, hipStreamWaitEvent(a->stream, a->ev_join, 0), "wait");
  if (!rc) {
    host_restore(a, hck[nseg]);
    rc = slab_from(a, tl_slot(a, nseg));
  }
  if (!rc) rc = tl_copy(a, false, cur, out, ld);
  if (!rc) rc = ahip(a, hipStreamSynchronize(a->stream), "sync");
  if (rc) {
    // nothing left running; then put the batch back at the window's end (the
    // recomputes moved its counters and fields), or refuse every later call
    const std::string msg = a->err;
    const bool synced = hipStreamSynchronize(a->astream) == hipSuccess &&
                        hipStreamSynchronize(a->stream) == hipSuccess;
    if (window_done) {
      host_restore(a, hck[nseg]);
      if (!synced || slab_from(a, tl_slot(a, nseg)) != GMC_OK || hipStreamSynchronize(a->stream) != hipSuccess)
        a->lost = true;
    }
    a->err = msg;
  }
  return rc;
}

}  // namespace

extern "C" {

int gmc_adv_ad_drive(gmc_adv *a, int64_t time_step0, const double *scale, int32_t nsteps, int32_t ckpt_every,
                     double *lam, int64_t ld) {
  return adv_adjoint(a, time_step0, scale, nsteps, ckpt_every, lam, nullptr, lam, nullptr, ld);
}

int gmc_adv_gradient(gmc_adv *a, int64_t time_step0, const double *scale, int32_t nsteps, int32_t ckpt_every,
                     const double *y, double *grad, double *cost, int64_t ld) {
  return adv_adjoint(a, time_step0, scale, nsteps, ckpt_every, nullptr, y, grad, cost, ld);
}

}  // extern "C"
