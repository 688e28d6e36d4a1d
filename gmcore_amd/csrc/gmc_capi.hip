// gmc_capi.hip -- device context and the C ABI declared in
// include/gmcore_hip.h.
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

#include "gmc_internal.hpp"

namespace gmc {

int set_error(gmc_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int check_hip(gmc_ctx *ctx, hipError_t e, const char *what) {
  if (e == hipSuccess) return GMC_OK;
  return set_error(ctx, e == hipErrorOutOfMemory ? GMC_ENOMEM : GMC_EDEVICE,
                   std::string(what) + ": " + hipGetErrorString(e));
}

// pc2 / wrfrk3 / euler on the operator-level kernels
// (time_schemes_mod.F90:102-118, 244-288)
static int ops_step_fb(gmc_ctx *c, int o, int s, int n, int t1, int t2, double dt) {
  int rc = ops_space_operators(c, o, s, n, t1, t2, dt, 1);
  if (!rc) rc = ops_update_state(c, t1, o, n, dt, 1);
  if (!rc) rc = ops_space_operators(c, o, s, n, t2, t1, dt, 2);
  if (!rc) rc = ops_update_state(c, t2, o, n, dt, 2);
  return rc;
}

int ops_integrate(gmc_ctx *c, int o, int n, double dt) {
  int rc;
  switch (c->cfg.time_scheme) {
  case GMC_TIME_EULER:
    rc = ops_step_fb(c, o, o, n, o, n, dt);
    break;
  case GMC_TIME_WRFRK3:
    rc = ops_step_fb(c, o, o, n, o, n, dt / 3.0);
    if (!rc) rc = ops_step_fb(c, o, n, 3, o, n, dt / 2.0);
    if (!rc) rc = ops_step_fb(c, o, 3, n, o, n, dt);
    break;
  default:
    rc = ops_step_fb(c, o, o, n, o, n, dt / 2.0);
    if (!rc) rc = ops_step_fb(c, o, n, 3, o, n, dt / 2.0);
    if (!rc) rc = ops_step_fb(c, o, 3, n, o, n, dt);
  }
  if (!rc && c->cfg.use_div_damp) rc = ops_div_damp(c, n);
  return rc;
}

static double *field_ptr(gmc_ctx *c, int level, int field) {
  if (level < 1 || level > 3) return nullptr;
  double **S = c->state[level], **T = c->tend[level];
  switch (field) {
  case GMC_U_LON: return S[S_U_LON];
  case GMC_V_LAT: return S[S_V_LAT];
  case GMC_GZ: return S[S_GZ];
  case GMC_U_F: return S[S_U_F];
  case GMC_V_F: return S[S_V_F];
  case GMC_GZ_F: return S[S_GZ_F];
  case GMC_GZS: return c->gzs;
  case GMC_DU: return T[T_DU];
  case GMC_DV: return T[T_DV];
  case GMC_DGZ: return T[T_DGZ];
  case GMC_PV: return S[S_PV];
  case GMC_KE: return S[S_KE];
  case GMC_DIV: return S[S_DIV];
  case GMC_M: return S[S_M];
  }
  return nullptr;
}

static int field_rows(const gmc_ctx *c, int field) {
  switch (field) {
  case GMC_V_LAT: case GMC_V_F: case GMC_DV: case GMC_PV: return c->grid.nhalf;
  default: return c->grid.nlat;
  }
}

static int upload_filter(gmc_ctx *c, const gmc_filter &f, FilterDev &d, int slot) {
  const int nlat = c->grid.nlat, nhalf = c->grid.nhalf;
  if (f.max_taps < 1 || !f.ngrid_lon || !f.ngrid_lat || !f.wgt_lon || !f.wgt_lat)
    return set_error(c, GMC_EINVAL, "filter table incomplete");
  for (int r = 0; r < nlat; r++)
    if (f.ngrid_lon[r] > f.max_taps || (f.ngrid_lon[r] > 1 && f.ngrid_lon[r] % 2 == 0))
      return set_error(c, GMC_EINVAL, "filter tap count out of range");
  for (int r = 0; r < nhalf; r++)
    if (f.ngrid_lat[r] > f.max_taps || (f.ngrid_lat[r] > 1 && f.ngrid_lat[r] % 2 == 0))
      return set_error(c, GMC_EINVAL, "filter tap count out of range");
  const size_t ni = sizeof(int32_t) * nlat * 2;
  const size_t nw = sizeof(double) * (size_t)nlat * f.max_taps * 2;
  void *mem = nullptr;
  int rc = check_hip(c, hipMalloc(&mem, ni + nw), "filter alloc");
  if (rc) return rc;
  c->d_filter_mem[slot] = mem;
  int32_t *ng = (int32_t *)mem;
  double *w = (double *)((char *)mem + ni);
  d.max_taps = f.max_taps;
  d.ngrid_lon = ng;
  d.ngrid_lat = ng + nlat;
  d.wgt_lon = w;
  d.wgt_lat = w + (size_t)nlat * f.max_taps;
  rc = check_hip(c, hipMemset(mem, 0, ni + nw), "filter memset");
  if (!rc) rc = check_hip(c, hipMemcpy((void *)d.ngrid_lon, f.ngrid_lon, sizeof(int32_t) * nlat, hipMemcpyHostToDevice), "filter copy");
  if (!rc) rc = check_hip(c, hipMemcpy((void *)d.ngrid_lat, f.ngrid_lat, sizeof(int32_t) * nhalf, hipMemcpyHostToDevice), "filter copy");
  if (!rc) rc = check_hip(c, hipMemcpy((void *)d.wgt_lon, f.wgt_lon, sizeof(double) * nlat * f.max_taps, hipMemcpyHostToDevice), "filter copy");
  if (!rc) rc = check_hip(c, hipMemcpy((void *)d.wgt_lat, f.wgt_lat, sizeof(double) * nhalf * f.max_taps, hipMemcpyHostToDevice), "filter copy");
  return rc;
}

hipError_t stage_copy(HostStage &s, hipStream_t stream, bool to_device, double *dev, int64_t pitch, double *host,
                      int64_t ld, int nlon, int rows) {
  const size_t n = (size_t)nlon * rows, row = sizeof(double) * nlon;
  if (s.cap < n) {
    stage_free(s);
    hipError_t e = hipHostMalloc((void **)&s.p, n * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
      s.p = nullptr;
      return e;
    }
    s.cap = n;
  }
  if (to_device) {
    for (int r = 0; r < rows; r++) std::memcpy(s.p + (size_t)r * nlon, host + r * ld, row);
    hipError_t e = hipMemcpy2DAsync(dev, sizeof(double) * pitch, s.p, row, row, rows, hipMemcpyHostToDevice, stream);
    return e == hipSuccess ? hipStreamSynchronize(stream) : e;
  }
  hipError_t e = hipMemcpy2DAsync(s.p, row, dev, sizeof(double) * pitch, row, rows, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e == hipSuccess)
    for (int r = 0; r < rows; r++) std::memcpy(host + r * ld, s.p + (size_t)r * nlon, row);
  return e;
}

void stage_free(HostStage &s) {
  if (s.p) (void)hipHostFree(s.p);
  s.p = nullptr;
  s.cap = 0;
}

}  // namespace gmc

using namespace gmc;

extern "C" {

int gmc_create(const gmc_mesh *mesh, const gmc_config *cfg, gmc_ctx **out) {
  if (!out) return GMC_EINVAL;
  *out = nullptr;
  if (!mesh || !cfg || !mesh->metrics) return GMC_EINVAL;
  if (mesh->nlon < 8 || mesh->nlon % 2 || mesh->nlat < 5) return GMC_EINVAL;
  if (cfg->coriolis_scheme != 1 && cfg->coriolis_scheme != 2) return GMC_EINVAL;
  if (cfg->upwind_order_pv != 1 && cfg->upwind_order_pv != 3) return GMC_EINVAL;
  if (cfg->use_div_damp && cfg->div_damp_order != 2 && cfg->div_damp_order != 4) return GMC_EINVAL;
  if (cfg->pv_scheme < GMC_PV_MIDPOINT || cfg->pv_scheme > GMC_PV_TVD) return GMC_EINVAL;
  if (cfg->pv_scheme == GMC_PV_TVD && (cfg->tvd_limiter < GMC_TVD_NONE || cfg->tvd_limiter > GMC_TVD_MC))
    return GMC_EINVAL;
  if (cfg->pv_scheme == GMC_PV_TVD && !(cfg->dt_dyn > 0)) return GMC_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return GMC_EDEVICE;
  gmc_ctx *c = new gmc_ctx();
  c->cfg = *cfg;
  c->dlon = mesh->dlon;
  Grid &g = c->grid;
  g.nlon = mesh->nlon;
  g.nlat = mesh->nlat;
  g.nhalf = mesh->nlat - 1;
  g.half = mesh->nlon / 2;
  g.pitch = (mesh->nlon + 31) / 32 * 32;
  int rc = check_hip(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream");
  const size_t fbytes = sizeof(double) * g.pitch * g.nlat;
  const size_t nf = 3 * (S_NFIELDS + T_NFIELDS) + 1;
  c->slab_bytes = fbytes * nf + sizeof(double) * (64 + (size_t)g.nlat * 8);
  if (!rc) rc = check_hip(c, hipMalloc((void **)&c->slab, c->slab_bytes), "field alloc");
  if (!rc) rc = check_hip(c, hipMemset(c->slab, 0, c->slab_bytes), "field memset");
  if (!rc) {
    char *p = (char *)c->slab;
    for (int l = 1; l <= 3; l++) {
      for (int f = 0; f < S_NFIELDS; f++, p += fbytes) c->state[l][f] = (double *)p;
      for (int f = 0; f < T_NFIELDS; f++, p += fbytes) c->tend[l][f] = (double *)p;
    }
    c->gzs = (double *)p; p += fbytes;
    c->d_pole = (double *)p; p += sizeof(double) * 64;
    c->d_diag = (double *)p;
  }
  if (!rc) rc = check_hip(c, hipMalloc((void **)&c->d_mesh, sizeof(double) * GMC_MESH_NFIELDS * g.nlat), "mesh alloc");
  if (!rc) rc = check_hip(c, hipMemcpy(c->d_mesh, mesh->metrics, sizeof(double) * GMC_MESH_NFIELDS * g.nlat,
                                       hipMemcpyHostToDevice), "mesh copy");
  if (!rc) {
    // row records of the metric table (zero records before the first and
    // after the last row): one wavefront reads its rows with a few wide loads.
    // Each record ends with the correctly rounded reciprocals of the divisor
    // fields (kRecipFields; 0 for a zero entry, which no formula divides by).
    std::vector<double> rows((size_t)(g.nlat + 2) * kRowRec, 0.0);
    for (int r = 0; r < g.nlat; r++) {
      double *rec = &rows[(size_t)(r + 1) * kRowRec];
      for (int f = 0; f < GMC_MESH_NFIELDS; f++) rec[f] = mesh->metrics[f * g.nlat + r];
      for (int k = 0; k < kNumRecip; k++) {
        const double d = rec[kRecipFields[k]];
        rec[GMC_MESH_NFIELDS + k] = d != 0.0 ? 1.0 / d : 0.0;
      }
    }
    rc = check_hip(c, hipMalloc((void **)&c->d_mrows, sizeof(double) * rows.size()), "mesh rows alloc");
    if (!rc) rc = check_hip(c, hipMemcpy(c->d_mrows, rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice),
                            "mesh rows copy");
  }
  if (!rc) rc = upload_filter(c, cfg->big_filter, c->big, 0);
  if (!rc) rc = upload_filter(c, cfg->small_filter1, c->small1, 1);
  if (!rc) rc = ops_scratch_init(c);
  if (rc) {
    std::fprintf(stderr, "gmc_create: %s\n", c->err.c_str());
    gmc_destroy(c);
    return rc;
  }
  *out = c;
  return GMC_OK;
}

int gmc_destroy(gmc_ctx *c) {
  if (!c) return GMC_EINVAL;
  if (c->stream) hip_note(hipStreamSynchronize(c->stream), "context sync");
  fused_release(c);
  hip_note(hipFree(c->slab), "field slab free");
  hip_note(hipFree(c->d_mesh), "mesh free");
  hip_note(hipFree(c->d_mrows), "mesh rows free");
  for (void *m : c->d_filter_mem) hip_note(hipFree(m), "filter table free");
  if (c->stream) hip_note(hipStreamDestroy(c->stream), "stream destroy");
  stage_free(c->stage);
  delete c;
  return GMC_OK;
}

const char *gmc_last_error(const gmc_ctx *c) { return c ? c->err.c_str() : "null context"; }

int gmc_upload(gmc_ctx *c, int32_t level, int32_t field, const double *host, int64_t ld) {
  if (!c || !host) return GMC_EINVAL;
  double *d = field_ptr(c, field == GMC_GZS ? 1 : level, field);
  if (!d || ld < c->grid.nlon) return set_error(c, GMC_EINVAL, "bad level/field/ld");
  return check_hip(c, stage_copy(c->stage, c->stream, true, d, c->grid.pitch, const_cast<double *>(host), ld,
                                 c->grid.nlon, field_rows(c, field)), "upload");
}

int gmc_download(gmc_ctx *c, int32_t level, int32_t field, double *host, int64_t ld) {
  if (!c || !host) return GMC_EINVAL;
  double *d = field_ptr(c, field == GMC_GZS ? 1 : level, field);
  if (!d || ld < c->grid.nlon) return set_error(c, GMC_EINVAL, "bad level/field/ld");
  return check_hip(c, stage_copy(c->stage, c->stream, false, d, c->grid.pitch, host, ld, c->grid.nlon,
                                 field_rows(c, field)), "download");
}

int gmc_prepare(gmc_ctx *c, int32_t level) {
  if (!c) return GMC_EINVAL;
  if (level < 1 || level > 3) return set_error(c, GMC_EINVAL, "level must be 1..3");
  int rc = ops_prepare(c, level);
  if (!rc) {
    c->derived_kept[level] = 1;
    c->old_level = level;
    c->new_level = level == 1 ? 2 : 1;
    c->prepared = 1;
  }
  return rc;
}

int gmc_space_operators(gmc_ctx *c, int32_t old_level, int32_t star_level, int32_t new_level, int32_t tend1,
                        int32_t tend2, double dt, int32_t pass) {
  if (!c) return GMC_EINVAL;
  for (int l : {old_level, star_level, new_level, tend1, tend2})
    if (l < 1 || l > 3) return set_error(c, GMC_EINVAL, "level must be 1..3");
  const int rc = ops_space_operators(c, old_level, star_level, new_level, tend1, tend2, dt, pass);
  if (!rc && pass != 2) c->derived_kept[star_level] = 1;   // operators_prepare_2 ran on the star state
  return rc;
}

int gmc_update_state(gmc_ctx *c, int32_t tend, int32_t old_level, int32_t new_level, double dt, int32_t what) {
  if (!c) return GMC_EINVAL;
  for (int l : {tend, old_level, new_level})
    if (l < 1 || l > 3) return set_error(c, GMC_EINVAL, "level must be 1..3");
  if (what < 1 || what > 3) return set_error(c, GMC_EINVAL, "what must be 1, 2 or 3");
  return ops_update_state(c, tend, old_level, new_level, dt, what);
}

int gmc_step(gmc_ctx *c, double dt, int32_t nsteps) {
  if (!c) return GMC_EINVAL;
  if (!c->prepared) return set_error(c, GMC_ESTATE, "gmc_prepare must run before gmc_step");
  if (nsteps < 0) return set_error(c, GMC_EINVAL, "nsteps < 0");
  constexpr int kStepBlock = 8;   // steps per graph of a long run (even: the levels end where they began)
  for (int n = 0; n < nsteps;) {
    const int block = nsteps - n >= kStepBlock ? kStepBlock : 1;
    int rc = fused_step_graph(c, c->old_level, c->new_level, dt, block);
    if (rc) return rc;
    for (int s = 0; s < block; s++) {
      derived_after_step(c, c->old_level, c->new_level);
      std::swap(c->old_level, c->new_level);
    }
    n += block;
  }
  return GMC_OK;
}

int gmc_time_integrate(gmc_ctx *c, int32_t old_level, int32_t new_level, double dt) {
  if (!c) return GMC_EINVAL;
  if (old_level < 1 || old_level > 2 || new_level < 1 || new_level > 2 || old_level == new_level)
    return set_error(c, GMC_EINVAL, "old and new must be the distinct levels 1 and 2");
  const int rc = fused_integrate(c, old_level, new_level, dt);
  if (!rc) derived_after_step(c, old_level, new_level);
  return rc;
}

int gmc_kernel_times(gmc_ctx *c, double dt, int32_t nsteps, double *out_ms, int32_t cap, int32_t *n) {
  if (!c || !out_ms || !n || cap < 1) return GMC_EINVAL;
  if (!c->prepared) return set_error(c, GMC_ESTATE, "gmc_prepare must run before gmc_kernel_times");
  if (nsteps < 1) return set_error(c, GMC_EINVAL, "nsteps < 1");
  int rc = fused_prepare(c);
  if (rc) return rc;
  const int stages = c->cfg.time_scheme == GMC_TIME_EULER ? 1 : 3;
  const int per_stage = 2 * kKernelClasses;   // a start and an end stamp per launch
  std::vector<hipEvent_t> ev((size_t)nsteps * stages * per_stage);
  for (auto &e : ev) {
    rc = check_hip(c, hipEventCreate(&e), "event create");
    if (rc) return rc;
  }
  c->prof_ev = ev.data();
  c->prof_cap = (int)ev.size();
  c->prof_n = 0;
  for (int s = 0; s < nsteps && !rc; s++) {
    rc = fused_integrate(c, c->old_level, c->new_level, dt);
    derived_after_step(c, c->old_level, c->new_level);
    std::swap(c->old_level, c->new_level);
  }
  c->prof_ev = nullptr;
  if (!rc) rc = check_hip(c, hipStreamSynchronize(c->stream), "timing sync");
  double acc[kKernelClasses] = {};
  int cnt = 0;
  for (int b = 0; !rc && b + per_stage <= c->prof_n; b += per_stage, cnt++)
    for (int k = 0; k < kKernelClasses; k++) {
      float ms = 0.f;
      rc = check_hip(c, hipEventElapsedTime(&ms, ev[b + 2 * k], ev[b + 2 * k + 1]), "event elapsed");
      acc[k] += ms;
    }
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  *n = std::min(cap, kKernelClasses);
  for (int k = 0; k < *n; k++) out_ms[k] = cnt ? acc[k] / cnt : 0.0;
  return GMC_OK;
}

const char *gmc_kernel_name(int32_t k) { return k >= 0 && k < kKernelClasses ? kKernelNames[k] : ""; }

int gmc_stream(gmc_ctx *c, void **stream) {
  if (!c || !stream) return GMC_EINVAL;
  *stream = (void *)c->stream;
  return GMC_OK;
}

int gmc_current_level(const gmc_ctx *c, int32_t *old_level) {
  if (!c || !old_level) return GMC_EINVAL;
  *old_level = c->old_level;
  return GMC_OK;
}

int gmc_diagnose(gmc_ctx *c, int32_t level, double out[4]) {
  if (!c || !out) return GMC_EINVAL;
  if (level < 1 || level > 3) return set_error(c, GMC_EINVAL, "level must be 1..3");
  return ops_diagnose(c, level, out);
}

int gmc_diagnose_dev(gmc_ctx *c, int32_t level, double *d_out) {
  if (!c || !d_out) return GMC_EINVAL;
  if (level < 1 || level > 3) return set_error(c, GMC_EINVAL, "level must be 1..3");
  return ops_diagnose_dev(c, level, d_out);
}

int gmc_synchronize(gmc_ctx *c) {
  if (!c) return GMC_EINVAL;
  return check_hip(c, hipStreamSynchronize(c->stream), "synchronize");
}

int gmc_field_ptr(gmc_ctx *c, int32_t level, int32_t field, double **dptr, int64_t *pitch) {
  if (!c || !dptr || !pitch) return GMC_EINVAL;
  double *d = field_ptr(c, field == GMC_GZS ? 1 : level, field);
  if (!d) return set_error(c, GMC_EINVAL, "bad level/field");
  *dptr = d;
  *pitch = c->grid.pitch;
  return GMC_OK;
}

int gmc_keep_derived(gmc_ctx *c, int32_t on) {
  if (!c) return GMC_EINVAL;
  if (on != 0 && on != 1) return set_error(c, GMC_EINVAL, "on must be 0 or 1");
  c->keep_derived = on;
  return GMC_OK;
}

int gmc_filter_plan(gmc_ctx *c, int32_t out[4]) {
  if (!c || !out) return GMC_EINVAL;
  return fused_plan_info(c, out);
}

}  // extern "C"
