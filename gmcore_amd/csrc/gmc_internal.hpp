// gmc_internal.hpp -- device context and shared device formulas of the
// gfx950 shallow-water dynamics library.
//
// Layout in HBM: every field is a dense row-major array [rows][pitch] of
// doubles, longitude fastest, no halo.  Periodic longitude neighbours are
// reached by modular column arithmetic; the cross-pole halo rows of the
// reference (parallel_mod.F90:529-563) by mirrored row indices with a
// half-circle column shift.  pitch is nlon rounded up to 32 doubles (256 B),
// so every row starts on a 256-byte boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/gmcore_hip.h"

namespace gmc {

constexpr double kEps = 2.220446049250313e-16;  // epsilon(1.0_r8)
constexpr int kWave = 64;

// state_type fields kept on the device (state_mod.F90:22-59)
enum StateField {
  S_U_LON, S_V_LAT, S_GZ, S_GZ_F, S_U_F, S_V_F, S_M, S_M_LON, S_M_LAT, S_M_VTX,
  S_MFX_LON, S_MFY_LAT, S_MFX_LAT, S_MFY_LON, S_U_LAT, S_V_LON, S_KE, S_DIV,
  S_VOR, S_PV, S_PV_LON, S_PV_LAT, S_DIV2, S_NFIELDS
};
// tend_type fields (tend_mod.F90:16-42)
enum TendField {
  T_DU, T_DV, T_DGZ, T_QHU, T_QHV, T_DKEDLON, T_DKEDLAT, T_DMFDLON, T_DMFDLAT,
  T_PGF_LON, T_PGF_LAT, T_NFIELDS
};

// a / d correctly rounded, from y = 1 / d correctly rounded: q0 = a y, the
// exact remainder e = a - d q0 (one fma), q = q0 + e y rounded once
// (Markstein's correction step, P. W. Markstein, IBM J. Res. Dev. 34 (1990)
// 111-119).  Gives the bits of the IEEE quotient for normal a, d and a / d;
// e == 0 returns q0, which keeps the sign of a zero numerator.  Three FP64
// operations instead of the scale / reciprocal / Newton / fixup sequence of a
// division, for denominators shared by a wavefront (metric terms of one row).
__device__ __forceinline__ double div_rn(double a, double d, double y) {
  const double q0 = a * y;
  const double e = __builtin_fma(-q0, d, a);
  return e == 0.0 ? q0 : __builtin_fma(e, y, q0);
}

// Metric fields whose reciprocals the row records also hold (slot order).
constexpr int kRecipFields[] = {GMC_AREA_VTX, GMC_AREA_CELL, GMC_DE_LON, GMC_DE_LAT};
constexpr int kNumRecip = 4;
constexpr int kRowRec = GMC_MESH_NFIELDS + kNumRecip;   // doubles per row record
__host__ __device__ constexpr int recip_slot(int f) {
  return f == GMC_AREA_VTX ? 0 : f == GMC_AREA_CELL ? 1 : f == GMC_DE_LON ? 2 : f == GMC_DE_LAT ? 3 : -1;
}

// Read-only view of the packed metric table: M(f, r).  M.div(x, f, r) is
// x / M(f, r) (the point formulas' divisions by a metric term).
struct Mesh {
  const double *m;
  int nlat;
  __device__ __forceinline__ double operator()(int f, int r) const { return m[f * nlat + r]; }
  __device__ __forceinline__ double div(double x, int f, int r) const { return x / (*this)(f, r); }
};

struct FilterDev {
  int max_taps;
  const int32_t *ngrid_lon, *ngrid_lat;
  const double *wgt_lon, *wgt_lat;
};

// Pinned host buffer for the host <-> device transfers of pitched fields:
// 2D copies between pageable host memory and pitched device rows go row by
// row, so a field passes through this buffer (packed rows) instead.
struct HostStage {
  double *p = nullptr;
  size_t cap = 0;   // doubles
};
// rows x nlon doubles between dev (row pitch `pitch`) and host (leading
// dimension ld), through s; returns once the copy is complete.
hipError_t stage_copy(HostStage &s, hipStream_t stream, bool to_device, double *dev, int64_t pitch, double *host,
                      int64_t ld, int nlon, int rows);
void stage_free(HostStage &s);

// Geometry shared by all kernels of a context.
struct Grid {
  int nlon, nlat, nhalf;
  int64_t pitch;
  int half;  // nlon / 2
};

// Workgroup b of a launch runs on XCD b % kNumXcd, and each XCD has its own
// L2.  xcd_item gives the item workgroup b should take so that every XCD
// works through one contiguous run of the items [0, total) (a launch of
// exactly total workgroups): XCD x takes [x q + min(x, m), ...), q = total /
// kNumXcd, m = total % kNumXcd, in order.  Items that share rows then share
// an L2.
constexpr int kNumXcd = 8;
__device__ __forceinline__ int xcd_item(int b, int total) {
  const int x = b % kNumXcd, s = b / kNumXcd, q = total / kNumXcd, m = total % kNumXcd;
  return x * q + (x < m ? x : m) + s;
}

// Column i+d, periodic.
__device__ __forceinline__ int wrap(int i, int n) {
  return i < 0 ? i + n : (i >= n ? i - n : i);
}

// Row index through the cross-pole mirror (halo_mod.F90:183-238 on one rank):
// full-lat row -k <- row k, row nlat-1+k <- nlat-1-k; half-lat row -k <- k-1,
// nhalf-1+k <- nhalf-k.  Mirrored rows take the column half a circle away.
__device__ __forceinline__ void mirror_full(int &r, int &i, const Grid &g) {
  if (r < 0) { r = -r; i = wrap(i + g.half, g.nlon); }
  else if (r > g.nlat - 1) { r = 2 * (g.nlat - 1) - r; i = wrap(i + g.half, g.nlon); }
}
__device__ __forceinline__ void mirror_half(int &r, int &i, const Grid &g) {
  if (r < 0) { r = -1 - r; i = wrap(i + g.half, g.nlon); }
  else if (r > g.nhalf - 1) { r = 2 * g.nhalf - 1 - r; i = wrap(i + g.half, g.nlon); }
}

// ---- point formulas (one per reference expression) ----------------------
// All use the source's left-to-right evaluation; the library is built with
// -ffp-contract=off so each product and sum is rounded separately.

// upwind1 / upwind3, upwind_mod.F90:14-39
__device__ __forceinline__ double upwind1(double dir, double wgt, double f1, double f2) {
  return 0.5 * (f2 + f1) + -0.5 * (f2 - f1) * wgt * dir;
}
__device__ __forceinline__ double upwind3(double dir, double wgt, double f1, double f2,
                                          double f3, double f4) {
  const double c31 = 7.0 / 12.0, c32 = -1.0 / 12.0, c33 = 1.0 / 12.0;
  return c31 * (f3 + f2) + c32 * (f4 + f1) + c33 * (f4 - f1 - 3 * (f3 - f2)) * wgt * dir;
}
}  // namespace gmc

// ---- device context ------------------------------------------------------
struct gmc_ctx {
  gmc::Grid grid;
  gmc_config cfg;            // host copy; filter pointers point at device tables
  double dlon;
  double *d_mesh = nullptr;  // GMC_MESH_NFIELDS * nlat
  double *d_mrows = nullptr; // row records: (nlat + 2) x kRowRec, record r + 1 = row r
  gmc::FilterDev big{}, small1{};
  void *d_filter_mem[8] = {};
  double *state[4][gmc::S_NFIELDS] = {};
  double *tend[4][gmc::T_NFIELDS] = {};
  double *gzs = nullptr;
  double *d_pole = nullptr;  // pole sums scratch (see gmc_ops.hip)
  double *d_diag = nullptr;  // diagnose partial sums
  double *scratch = nullptr; // this device's static scratch of gmc_diagnose_dev
  double *slab = nullptr;    // one allocation holding every field
  size_t slab_bytes = 0;
  hipStream_t stream = nullptr;
  gmc::HostStage stage;
  int old_level = 1, new_level = 2;
  int prepared = 0;
  // derived fields: with keep_derived, every fused stage also stores the
  // star state's forward-pass fields in the star level (as the reference's
  // operators leave them in state(star)); derived_kept[l] says level l holds
  // the fields its last use as a star state (or gmc_prepare) left there
  int keep_derived = 0;
  int derived_kept[4] = {0, 0, 0, 0};
  void *fused_plan = nullptr;  // work lists of the fused stage (gmc_fused.hip)
  // graphs of one fused step, keyed by (old, new, dt, keep_derived)
  struct StepGraph { int o, n; double dt; int keep, steps; hipGraphExec_t exec; };
  std::vector<StepGraph> graphs;
  // kernel timing: when non-null, each fused_stage launch takes the next two
  // events as its start and end stamps
  hipEvent_t *prof_ev = nullptr;
  int prof_n = 0, prof_cap = 0;
  std::string err;
};

namespace gmc {
// A HIP call on a release path, whose failure the caller cannot return
// (destroy, graph and buffer release): reported on stderr.
inline void hip_note(hipError_t e, const char *what) {
  if (e != hipSuccess) std::fprintf(stderr, "gmcore_hip: %s: %s\n", what, hipGetErrorString(e));
}
int set_error(gmc_ctx *ctx, int code, const std::string &msg);
int check_hip(gmc_ctx *ctx, hipError_t e, const char *what);
// operator-level path (gmc_ops.hip)
int ops_prepare(gmc_ctx *ctx, int level);
int ops_space_operators(gmc_ctx *ctx, int old_l, int star_l, int new_l, int t1, int t2, double dt,
                        int pass);
int ops_update_state(gmc_ctx *ctx, int tend, int old_l, int new_l, double dt, int what);
int ops_div_damp(gmc_ctx *ctx, int level);
// calc_div of state `level` (its filtered winds) into its div field
int ops_calc_div(gmc_ctx *ctx, int level);
int ops_diagnose(gmc_ctx *ctx, int level, double out[4]);
// resolve (once per device) the address of gmc_diagnose_dev's static scratch
int ops_scratch_init(gmc_ctx *ctx);
int ops_diagnose_dev(gmc_ctx *ctx, int level, double *d_out);
int ops_integrate(gmc_ctx *ctx, int old_l, int new_l, double dt);
void ops_filter(gmc_ctx *ctx, int small, int half, const double *x, double *y);
int ops_stage_tape(gmc_ctx *ctx, int old_l, int star_l, int new_l, double dt);
// fused stage path (gmc_fused.hip)
int fused_integrate(gmc_ctx *ctx, int old_l, int new_l, double dt);
void fused_release(gmc_ctx *ctx);
int fused_prepare(gmc_ctx *ctx);
int fused_plan_info(gmc_ctx *ctx, int32_t out[4]);
// host bookkeeping of derived_kept after one time step old -> new of the
// configured scheme (the star levels are old, new and 3; old alone for euler)
void derived_after_step(gmc_ctx *ctx, int old_l, int new_l);
// a path that steps without keeping derived fields (ensemble, TL/AD)
void derived_forget(gmc_ctx *ctx);
int fused_step_graph(gmc_ctx *ctx, int old_l, int new_l, double dt, int steps = 1);
// Destinations of the intermediates a taping stage writes for the adjoint:
// the fields operators_prepare forms from the star state (mfx_lat and
// mfy_lon only where not null); when u_f is not null the star state's u_f
// and v_f; when dlon[0] is not null the derivatives of the upwind pv on the
// pv_lon and pv_lat edges (dt, dn, df[0..3] of upwind_deriv, six fields
// each).
struct TapeOut {
  double *m_lon, *m_lat, *m_vtx, *mfx, *mfy, *mfx_lat, *mfy_lon, *u_lat, *v_lon, *pv, *pv_lon, *pv_lat;
  double *u_f, *v_f;
  double *dlon[6], *dlat[6];
};
// one fused stage that also leaves the star state's intermediates (the
// fields ops_stage_tape leaves for the adjoint) in its state fields, or in
// *dst when dst is given
int fused_stage_tape(gmc_ctx *ctx, int old_l, int star_l, int new_l, double dt, const TapeOut *dst = nullptr);
constexpr int kKernelClasses = 3;  // k_tend, k_gzf, k_uvf
extern const char *const kKernelNames[kKernelClasses];
}  // namespace gmc
