// gmc_baro.hip -- hydrostatic column operators of the baroclinic dynamics
// (include/gmcore_baro.h).
//
// Every kernel gives one thread one column (i, r) and walks the levels from
// the model top, so the loads and stores of a level are coalesced along the
// latitude row and the level recurrences (the geopotential's sums, the
// vertical mass flux) stay in the thread that owns the column.  Fields are
// [levels][rows][pitch] arrays in one device slab; V-point fields have
// nlat - 1 rows per level.  The arithmetic is the reference's, operation by
// operation and in its order (IEEE division, no contraction); pow and log are
// the device library's, within a few ulp of the host's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gmc_internal.hpp"
#include "gmc_point.hpp"
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
    {1, false}, {1, true}};                              // u_lon, v_lat

// calc_ph (operators_mod.F90:136-168): the half-level pressures of the
// hybrid coordinate (hybrid_coord_mod.F90:155-162), their Exner powers, and
// the full levels' mean of the two faces
__global__ void __launch_bounds__(NB) k_calc_ph(BaroGrid g, const double *__restrict__ hyai,
                                                const double *__restrict__ hybi, double p0, double ptop,
                                                double rd_o_cpd, const double *__restrict__ phs, double *ph_lev,
                                                double *ph_exn_lev, double *ph) {
  const int i = blockIdx.x * NB + threadIdx.x, r = blockIdx.y;
  if (i >= g.nlon) return;
  const int64_t o = r * g.pitch + i;
  const double ps = phs[o];
  double above = 0.0;
  for (int k = 0; k <= g.nlev; k++) {
    const double pl = hyai[k] * (p0 - ptop) + hybi[k] * (ps - ptop) + ptop;
    ph_lev[k * g.sc + o] = pl;
    ph_exn_lev[k * g.sc + o] = pow(pl, rd_o_cpd);
    if (k > 0) ph[(k - 1) * g.sc + o] = 0.5 * (above + pl);
    above = pl;
  }
}

// calc_t (operators_mod.F90:170-190) with temperature() (formula_mod.F90:24-31)
__global__ void __launch_bounds__(NB) k_calc_t(BaroGrid g, double p0, double rd_o_cpd, const double *__restrict__ pt,
                                               const double *__restrict__ ph, double *t) {
  const int i = blockIdx.x * NB + threadIdx.x, r = blockIdx.y;
  if (i >= g.nlon) return;
  const int64_t o = r * g.pitch + i;
  for (int k = 0; k < g.nlev; k++) t[k * g.sc + o] = pt[k * g.sc + o] * pow(ph[k * g.sc + o] / p0, rd_o_cpd);
}

// calc_gz_lev (operators_mod.F90:433-469).  Each half level's sum runs from
// 0 over the levels below it in ascending order, as the reference's inner
// loop does (so no running suffix sum: that would change the bits).  The
// layer terms Rd t log(ph_lev(l+1) / ph_lev(l)) are formed once into the gz
// array, which the full-level means overwrite at the end.
__global__ void __launch_bounds__(NB) k_calc_gz_lev(BaroGrid g, double rd, const double *__restrict__ gzs,
                                                    const double *__restrict__ t,
                                                    const double *__restrict__ ph_lev, double *gz_lev, double *gz) {
  const int i = blockIdx.x * NB + threadIdx.x, r = blockIdx.y;
  if (i >= g.nlon) return;
  const int64_t o = r * g.pitch + i;
  const int nl = g.nlev;
  for (int l = 0; l < nl; l++)
    gz[l * g.sc + o] = rd * t[l * g.sc + o] * log(ph_lev[(l + 1) * g.sc + o] / ph_lev[l * g.sc + o]);
  const double s = gzs[o];
  for (int k = 0; k < nl; k++) {
    double dgz = 0.0;
    for (int l = k; l < nl; l++) dgz = dgz + gz[l * g.sc + o];
    gz_lev[k * g.sc + o] = s + dgz;
  }
  for (int k = 0; k < nl; k++) gz[k * g.sc + o] = 0.5 * (gz_lev[k * g.sc + o] + gz_lev[(k + 1) * g.sc + o]);
}

// calc_dphsdt (operators_mod.F90:1120-1143)
__global__ void __launch_bounds__(NB) k_calc_dphsdt(BaroGrid g, const double *__restrict__ dmfdlon,
                                                    const double *__restrict__ dmfdlat, double *dphs) {
  const int i = blockIdx.x * NB + threadIdx.x, r = blockIdx.y;
  if (i >= g.nlon) return;
  const int64_t o = r * g.pitch + i;
  double d = 0.0;
  for (int k = 0; k < g.nlev; k++) d = d - dmfdlon[k * g.sc + o] - dmfdlat[k * g.sc + o];
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
  for (int k = 1; k < g.nlev; k++) {
    mf = mf + dmfdlon[(k - 1) * g.sc + o] + dmfdlat[(k - 1) * g.sc + o];
    we_lev[k * g.sc + o] = -(hybi[k] * dp) - mf;
  }
  we_lev[g.nlev * g.sc + o] = 0.0;
}

// interp_lev_edge_to_lev_lon_edge / _lat_edge with no wind
// (interp_mod.F90:360-369, :425-434), every half level: blockIdx.y < nlat - 2
// are the U rows 1 .. nlat-2, the rest the V rows 0 .. nlat-2
__global__ void __launch_bounds__(NB) k_we_edges(BaroGrid g, Mesh M, const double *__restrict__ we_lev,
                                                 double *we_lon, double *we_lat) {
  const int i = blockIdx.x * NB + threadIdx.x;
  if (i >= g.nlon) return;
  const int y = blockIdx.y;
  if (y < g.nlat - 2) {
    const int r = y + 1, ie = i + 1 == g.nlon ? 0 : i + 1;
    const double aw = M(GMC_AREA_LON_WEST, r), ae = M(GMC_AREA_LON_EAST, r), a = M(GMC_AREA_LON, r);
    const int64_t o = r * g.pitch;
    for (int k = 0; k <= g.nlev; k++)
      we_lon[k * g.sc + o + i] = (aw * we_lev[k * g.sc + o + i] + ae * we_lev[k * g.sc + o + ie]) / a;
  } else {
    const int r = y - (g.nlat - 2);
    const double an = M(GMC_AREA_LAT_NORTH, r), as = M(GMC_AREA_LAT_SOUTH, r), a = M(GMC_AREA_LAT, r);
    const int64_t o = r * g.pitch + i;
    for (int k = 0; k <= g.nlev; k++)
      we_lat[k * g.sv + o] = (an * we_lev[k * g.sc + o + g.pitch] + as * we_lev[k * g.sc + o]) / a;
  }
}

// calc_wedudlev_wedvdlev (operators_mod.F90:1145-1216), one column of one
// point kind: x the wind (u_f or v_f), we its half-level mass flux, m its
// full-level mass; the top and bottom levels take their one face
__device__ __forceinline__ void wedx_column(int nlev, int64_t s, const double *__restrict__ x,
                                            const double *__restrict__ we, const double *__restrict__ m,
                                            double *out) {
  double xm = 0.0, x0 = x[0], xp = x[s];
  out[0] = (we[s] * (xp - x0)) / m[0] / 2.0;
  for (int k = 1; k < nlev - 1; k++) {
    xm = x0;
    x0 = xp;
    xp = x[(k + 1) * s];
    out[k * s] = (we[(k + 1) * s] * (xp - x0) + we[k * s] * (x0 - xm)) / m[k * s] / 2.0;
  }
  const int k = nlev - 1;
  out[k * s] = (we[k * s] * (x[k * s] - x[(k - 1) * s])) / m[k * s] / 2.0;
}

__global__ void __launch_bounds__(NB) k_wedudlev(BaroGrid g, const double *__restrict__ u,
                                                 const double *__restrict__ we_lon, const double *__restrict__ m_lon,
                                                 double *wedudlev, const double *__restrict__ v,
                                                 const double *__restrict__ we_lat,
                                                 const double *__restrict__ m_lat, double *wedvdlev) {
  const int i = blockIdx.x * NB + threadIdx.x;
  if (i >= g.nlon) return;
  const int y = blockIdx.y;
  if (y < g.nlat - 2) {
    const int64_t o = (y + 1) * g.pitch + i;   // U rows 1 .. nlat-2
    wedx_column(g.nlev, g.sc, u + o, we_lon + o, m_lon + o, wedudlev + o);
  } else {
    const int64_t o = (y - (g.nlat - 2)) * g.pitch + i;   // V rows 0 .. nlat-2
    wedx_column(g.nlev, g.sv, v + o, we_lat + o, m_lat + o, wedvdlev + o);
  }
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

__global__ void __launch_bounds__(NB) k_pgf_lin97(BaroGrid g, Mesh M, const double *__restrict__ exn,
                                                  const double *__restrict__ gz, double *pgf_lon, double *pgf_lat) {
  const int i = blockIdx.x * NB + threadIdx.x;
  if (i >= g.nlon) return;
  const int y = blockIdx.y;
  const int64_t s = g.sc;
  if (y < g.nlat - 2) {
    const int r = y + 1, ie = i + 1 == g.nlon ? 0 : i + 1;   // U rows 1 .. nlat-2
    const double de = M(GMC_DE_LON, r);
    const int64_t a = r * g.pitch + i, b = r * g.pitch + ie;
    for (int k = 0; k < g.nlev; k++)
      pgf_lon[k * s + a] = lin97(exn[k * s + a], exn[(k + 1) * s + a], exn[k * s + b], exn[(k + 1) * s + b],
                                 gz[k * s + a], gz[(k + 1) * s + a], gz[k * s + b], gz[(k + 1) * s + b], de);
  } else {
    const int r = y - (g.nlat - 2);   // V rows 0 .. nlat-2
    const double de = M(GMC_DE_LAT, r);
    const int64_t a = r * g.pitch + i, b = a + g.pitch, o = r * g.pitch + i;
    for (int k = 0; k < g.nlev; k++)
      pgf_lat[k * g.sv + o] = lin97(exn[k * s + a], exn[(k + 1) * s + a], exn[k * s + b], exn[(k + 1) * s + b],
                                    gz[k * s + a], gz[(k + 1) * s + a], gz[k * s + b], gz[(k + 1) * s + b], de);
  }
}

// the vertical part of calc_grad_ptf (operators_mod.F90:1083-1101) with
// vflx_ppm_cell (ffsl_mod.F90:883-941).  Work arrays of the column: ext, pt
// on levels -2 .. nlev+1 (ext level e = k + 2, stride sc, nlev + 4 levels);
// ml, dm, m6, the PPM coefficients of the column's cells.  flag[0] is set
// where a Courant number reaches outside the column (the index is then
// clamped so that every access stays in the arrays).
__global__ void __launch_bounds__(NB) k_vflx_pt(BaroGrid g, int lim, const double *__restrict__ pt,
                                                const double *__restrict__ we, const double *__restrict__ cflz,
                                                double *ext, double *ml, double *dm, double *m6, double *ptf,
                                                double *dptf, int *flag) {
  const int i = blockIdx.x * NB + threadIdx.x, r = blockIdx.y;
  if (i >= g.nlon) return;
  const int64_t o = r * g.pitch + i, s = g.sc;
  const int n = g.nlev;
  for (int k = 0; k < n; k++) ext[(k + 2) * s + o] = pt[k * s + o];
  ext[1 * s + o] = 2 * ext[2 * s + o] - ext[3 * s + o];
  ext[o] = 2 * ext[1 * s + o] - ext[2 * s + o];
  ext[(n + 2) * s + o] = 2 * ext[(n + 1) * s + o] - ext[n * s + o];
  ext[(n + 3) * s + o] = 2 * ext[(n + 2) * s + o] - ext[(n + 1) * s + o];
  for (int k = 0; k < n; k++) {
    const double *e = ext + (k + 2) * s + o;
    const Ppm p = ppm(lim, e[-2 * s], e[-s], e[0], e[s], e[2 * s]);
    ml[k * s + o] = p.fl;
    dm[k * s + o] = p.df;
    m6[k * s + o] = p.f6;
  }
  bool bad = false;
  for (int k = 1; k < n; k++) {   // interior half levels
    const double c = cflz[k * s + o];
    int ci = (int)c;
    const double cf = c - ci;
    ci = ci > n + 4 ? n + 4 : (ci < -n - 4 ? -n - 4 : ci);   // beyond the column either way: flagged below
    double f = 0.0;
    if (c != 0) {
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
      for (int j = j0; j <= j1; j++) sum = sum + ext[(j + 2) * s + o];
      const double s1 = down ? 1 - cf : 0.0, s2 = down ? 1.0 : -cf;
      const double ds1 = s2 - s1, ds2 = s2 * s2 - s1 * s1, ds3 = s2 * s2 * s2 - s1 * s1 * s1;
      const double w = we[k * s + o];
      f = (down ? w : -w) *
          (sum + ml[ku * s + o] * ds1 + 0.5 * dm[ku * s + o] * ds2 + m6[ku * s + o] * (ds2 / 2.0 - ds3 / 3.0)) / c;
    }
    ptf[k * s + o] = f;
  }
  for (int k = 0; k < n; k++) dptf[k * s + o] = ptf[(k + 1) * s + o] - ptf[k * s + o];
  if (bad) flag[0] = 1;
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
  double *w_ext = nullptr, *w_ml = nullptr, *w_dm = nullptr, *w_m6 = nullptr;   // calc_ptf_lev work
  int *d_flag = nullptr;
  void *d_filter = nullptr;          // small_filter2's table on the device
  size_t filter_bytes = 0;
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
  n += (size_t)(nlev + 4 + 3 * nlev) * g.sc + 4;   // column work of calc_ptf_lev, the flag
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
  b->w_ext = p;
  p += (size_t)(nlev + 4) * g.sc;
  b->w_ml = p;
  p += (size_t)nlev * g.sc;
  b->w_dm = p;
  p += (size_t)nlev * g.sc;
  b->w_m6 = p;
  p += (size_t)nlev * g.sc;
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
  hipLaunchKernelGGL(k_calc_ph, columns(g, g.nlat), dim3(NB), 0, b->stream, g, b->d_hyai, b->d_hybi, b->vp0, b->ptop,
                     b->rd_o_cpd, b->f[GMB_PHS], b->f[GMB_PH_LEV], b->f[GMB_PH_EXN_LEV], b->f[GMB_PH]);
  return launched(b, "calc_ph");
}

int gmc_baro_calc_t(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(k_calc_t, columns(g, g.nlat), dim3(NB), 0, b->stream, g, b->tp0, b->rd_o_cpd, b->f[GMB_PT],
                     b->f[GMB_PH], b->f[GMB_T]);
  return launched(b, "calc_t");
}

int gmc_baro_calc_gz_lev(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(k_calc_gz_lev, columns(g, g.nlat), dim3(NB), 0, b->stream, g, b->Rd, b->f[GMB_GZS], b->f[GMB_T],
                     b->f[GMB_PH_LEV], b->f[GMB_GZ_LEV], b->f[GMB_GZ]);
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
  hipLaunchKernelGGL(k_we_edges, columns(g, 2 * g.nlat - 3), dim3(NB), 0, b->stream, g, Mesh{b->d_mesh, g.nlat},
                     b->f[GMB_WE_LEV], b->f[GMB_WE_LEV_LON], b->f[GMB_WE_LEV_LAT]);
  return launched(b, "calc_we_lev");
}

int gmc_baro_calc_wedudlev_wedvdlev(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(k_wedudlev, columns(g, 2 * g.nlat - 3), dim3(NB), 0, b->stream, g, b->f[GMB_U_F],
                     b->f[GMB_WE_LEV_LON], b->f[GMB_M_LON], b->f[GMB_WEDUDLEV], b->f[GMB_V_F], b->f[GMB_WE_LEV_LAT],
                     b->f[GMB_M_LAT], b->f[GMB_WEDVDLEV]);
  return launched(b, "calc_wedudlev_wedvdlev");
}

int gmc_baro_pgf_lin97(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  const BaroGrid &g = b->g;
  hipLaunchKernelGGL(k_pgf_lin97, columns(g, 2 * g.nlat - 3), dim3(NB), 0, b->stream, g, Mesh{b->d_mesh, g.nlat},
                     b->f[GMB_PH_EXN_LEV], b->f[GMB_GZ_LEV], b->f[GMB_PGF_LON], b->f[GMB_PGF_LAT]);
  return launched(b, "pgf_lin97");
}

int gmc_baro_calc_ptf_lev(gmc_baro *b, int32_t limiter) {
  if (!b) return GMC_EINVAL;
  if (limiter < GMC_LIMITER_MONO || limiter > GMC_LIMITER_NONE) return berr(b, GMC_EINVAL, "bad limiter");
  const BaroGrid &g = b->g;
  int rc = bhip(b, hipMemsetAsync(b->d_flag, 0, sizeof(int), b->stream), "flag clear");
  if (rc) return rc;
  hipLaunchKernelGGL(k_vflx_pt, columns(g, g.nlat), dim3(NB), 0, b->stream, g, (int)limiter, b->f[GMB_PT],
                     b->f[GMB_WE_LEV], b->f[GMB_CFLZ], b->w_ext, b->w_ml, b->w_dm, b->w_m6, b->f[GMB_PTF_LEV],
                     b->f[GMB_DPTFDLEV], b->d_flag);
  rc = launched(b, "calc_ptf_lev");
  int h = 0;
  if (!rc) rc = bhip(b, hipMemcpyAsync(&h, b->d_flag, sizeof(int), hipMemcpyDeviceToHost, b->stream), "flag");
  if (!rc) rc = bhip(b, hipStreamSynchronize(b->stream), "sync");
  if (!rc && h) rc = berr(b, GMC_ENUMERIC, "cflz reaches outside the column");
  return rc;
}

int gmc_baro_small_filter2(gmc_baro *b, const gmc_filter *f) {
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
  const size_t ni = sizeof(int32_t) * (nlat + nhalf), nw = sizeof(double) * (size_t)f->max_taps * (nlat + nhalf);
  const size_t need = (ni + 255) / 256 * 256 + nw;
  int rc = GMC_OK;
  if (need > b->filter_bytes) {
    if (b->d_filter) {
      rc = bhip(b, hipStreamSynchronize(b->stream), "sync");
      if (!rc) rc = bhip(b, hipFree(b->d_filter), "filter free");
      b->d_filter = nullptr;
      b->filter_bytes = 0;
    }
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
  const size_t lds = sizeof(double) * (g.nlon + 2 * (size_t)hmax);
  if (lds > 65536) return berr(b, GMC_EINVAL, "filter row does not fit in LDS");
  hipLaunchKernelGGL(k_filter_levels, dim3(nlat, g.nlev), dim3(NB), lds, b->stream, g, nlat, g.sc, ng, w,
                     f->max_taps, b->f[GMB_U_LON]);
  hipLaunchKernelGGL(k_filter_levels, dim3(nhalf, g.nlev), dim3(NB), lds, b->stream, g, nhalf, g.sv, ng + nlat,
                     (const double *)((const char *)w + wl), f->max_taps, b->f[GMB_V_LAT]);
  return launched(b, "small_filter2");
}

int gmc_baro_synchronize(gmc_baro *b) {
  if (!b) return GMC_EINVAL;
  return bhip(b, hipStreamSynchronize(b->stream), "synchronize");
}

int gmc_baro_stream(gmc_baro *b, void **stream) {
  if (!b || !stream) return GMC_EINVAL;
  *stream = (void *)b->stream;
  return GMC_OK;
}

}  // extern "C"
