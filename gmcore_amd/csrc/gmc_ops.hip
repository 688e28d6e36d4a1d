// gmc_ops.hip -- operator-level kernels of the shallow-water path.
//
// One kernel per reference operator (src/dynamics/operators_mod.F90,
// pgf/pgf_swm_mod.F90, interp_mod.F90, time_schemes_mod.F90:update_state),
// with every intermediate of state_type/tend_type materialised in HBM.  They
// back the per-pass entry points (gmc_space_operators, gmc_update_state,
// gmc_prepare, gmc_diagnose).  gmc_step uses the fused stage kernels of
// gmc_fused.hip, which call the same point formulas.
#include <mutex>

#include "gmc_internal.hpp"
#include "gmc_point.hpp"

namespace gmc {

namespace {

constexpr int BX = 64, BY = 4;

dim3 grid2d(const Grid &g, int rows) { return dim3((g.nlon + BX - 1) / BX, (rows + BY - 1) / BY); }

// ---------------------------------------------------------------------------
// calc_m SW branch + average_cell_to_lon_edge / _lat_edge / interp_cell_to_vtx
// (operators_mod.F90:527-541, interp_mod.F90:233-275, 438-463)
__global__ void k_calc_m(Grid g, Mesh M, double grav, const double *__restrict__ gz,
                         const double *__restrict__ gzs, double *m, double *m_lon, double *m_lat,
                         double *m_vtx) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch;
  const int ip = wrap(i + 1, g.nlon);
  auto mm = [&](int rr, int ii) { return m_of(gz[rr * p + ii], gzs[rr * p + ii], grav); };
  m[r * p + i] = mm(r, i);
  if (r >= 1 && r <= g.nlat - 2) m_lon[r * p + i] = (mm(r, i) + mm(r, ip)) * 0.5;
  if (r <= g.nhalf - 1) {
    m_lat[r * p + i] = (mm(r, i) + mm(r + 1, i)) * 0.5;
    m_vtx[r * p + i] = m_vtx_of(mm(r, i), mm(r, ip), mm(r + 1, i), mm(r + 1, ip), M, r);
  }
}

// calc_mf (operators_mod.F90:545-610).  mfx_lon is zero on the pole rows,
// which the reference never writes.
__global__ void k_calc_mf(Grid g, const double *__restrict__ m_lon, const double *__restrict__ m_lat,
                          const double *__restrict__ u_f, const double *__restrict__ v_f, Mesh M,
                          double *mfx_lon, double *mfy_lat, double *mfx_lat, double *mfy_lon,
                          double *u_lat, double *v_lon) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch;
  const int im = wrap(i - 1, g.nlon), ip = wrap(i + 1, g.nlon);
  auto mfx = [&](int rr, int ii) {
    return (rr == 0 || rr == g.nlat - 1) ? 0.0 : m_lon[rr * p + ii] * u_f[rr * p + ii];
  };
  auto mfy = [&](int rr, int ii) { return m_lat[rr * p + ii] * v_f[rr * p + ii]; };
  if (r >= 1 && r <= g.nlat - 2) {
    mfx_lon[r * p + i] = mfx(r, i);
    const double v = tangent_sum(M(GMC_FULL_TANGENT_WGT1, r), M(GMC_FULL_TANGENT_WGT2, r),
                                 mfy(r - 1, i), mfy(r - 1, ip), mfy(r, i), mfy(r, ip));
    mfy_lon[r * p + i] = v;
    v_lon[r * p + i] = v / m_lon[r * p + i];
  }
  if (r <= g.nhalf - 1) {
    mfy_lat[r * p + i] = mfy(r, i);
    const double v = tangent_sum(M(GMC_HALF_TANGENT_WGT1, r), M(GMC_HALF_TANGENT_WGT2, r),
                                 mfx(r, im), mfx(r, i), mfx(r + 1, im), mfx(r + 1, i));
    mfx_lat[r * p + i] = v;
    u_lat[r * p + i] = v / m_lat[r * p + i];
  }
}

// Pole reductions of one state (zonal_sum call sites in calc_ke :320-349,
// calc_div :378-407, calc_vor :640-674, calc_grad_mf :974-1003).  Block 0 is
// the south pole, block 1 the north pole; one wavefront each.
// pole[b*4 + 0] = ke, +1 = div, +2 = dmfdlat, +3 = vor (vertex row)
__global__ void k_pole_sums(Grid g, Mesh M, const double *__restrict__ v_f,
                            const double *__restrict__ mfy_lat, const double *__restrict__ u_lat,
                            double *pole) {
  const int north = blockIdx.x;
  const int lane = threadIdx.x;
  const int vr = north ? g.nhalf - 1 : 0;       // V row next to the pole
  const int cr = north ? g.nlat - 1 : 0;        // pole cell row
  const int64_t p = g.pitch;
  const double *v = v_f + vr * p, *f = mfy_lat + vr * p, *ul = u_lat + vr * p;
  const double le = M(GMC_LE_LAT, vr), ac = M(GMC_AREA_CELL, cr);
  const double nlon = (double)g.nlon;
  const double ske = wave_zonal_sum(g.nlon, lane, [&](int i) { return v[i] * v[i]; });
  const double sdv = wave_zonal_sum(g.nlon, lane, [&](int i) { return north ? -v[i] : v[i]; });
  const double smf = wave_zonal_sum(g.nlon, lane, [&](int i) { return north ? -f[i] : f[i]; });
  const double svo = wave_zonal_sum(g.nlon, lane, [&](int i) { return north ? ul[i] * le : -ul[i] * le; });
  if (lane == 0) {
    pole[north * 4 + 0] = ske / nlon;
    pole[north * 4 + 1] = sdv * le / nlon / ac;
    pole[north * 4 + 2] = smf * le / nlon / ac;
    pole[north * 4 + 3] = svo / nlon / ac;
  }
}

// calc_ke (ke_scheme 1), operators_mod.F90:234-258 + poles
__global__ void k_calc_ke(Grid g, Mesh M, const double *__restrict__ u, const double *__restrict__ v,
                          const double *__restrict__ pole, double *ke) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch;
  if (r == 0 || r == g.nlat - 1) { ke[r * p + i] = pole[(r == 0 ? 0 : 1) * 4 + 0]; return; }
  const int im = wrap(i - 1, g.nlon);
  ke[r * p + i] = ke_of(u[r * p + im], u[r * p + i], v[(r - 1) * p + i], v[r * p + i], M, r);
}

// calc_div, operators_mod.F90:354-412
__global__ void k_calc_div(Grid g, Mesh M, const double *__restrict__ u, const double *__restrict__ v,
                           const double *__restrict__ pole, double *div) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch;
  if (r == 0 || r == g.nlat - 1) { div[r * p + i] = pole[(r == 0 ? 0 : 1) * 4 + 1]; return; }
  const int im = wrap(i - 1, g.nlon);
  div[r * p + i] = div_of(u[r * p + i], u[r * p + im], v[r * p + i], v[(r - 1) * p + i], M, r);
}

// calc_vor + calc_pv, operators_mod.F90:612-703
__global__ void k_calc_pv(Grid g, Mesh M, int pole_stokes, const double *__restrict__ u,
                          const double *__restrict__ v, const double *__restrict__ m_vtx,
                          const double *__restrict__ pole, double *vor, double *pv) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nhalf) return;
  const int64_t p = g.pitch;
  const int ip = wrap(i + 1, g.nlon);
  double w = vor_of(u[r * p + i], u[(r + 1) * p + i], v[r * p + ip], v[r * p + i], M, r);
  if (pole_stokes && r == 0) w = pole[0 * 4 + 3];
  else if (pole_stokes && r == g.nhalf - 1) w = pole[1 * 4 + 3];
  vor[r * p + i] = w;
  pv[r * p + i] = pv_of(w, M, r, m_vtx[r * p + i]);
}

// interp_pv_upwind / interp_pv_midpoint / interp_pv_tvd, operators_mod.F90:705-832
// (order: upwind_order_pv, or the tvd limiter; dt: the stage dt)
__global__ void k_interp_pv(Grid g, Mesh M, int scheme, int order, double wgt, double dt,
                            const double *__restrict__ pv, const double *__restrict__ u_f,
                            const double *__restrict__ v_f, const double *__restrict__ u_lat,
                            const double *__restrict__ v_lon, double *pv_lon, double *pv_lat) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch;
  auto PV = [&](int rr, int ii) { mirror_half(rr, ii, g); return pv[rr * p + ii]; };
  if (r >= 1 && r <= g.nlat - 2) {
    const double f1 = PV(r - 2, i), f2 = PV(r - 1, i), f3 = PV(r, i), f4 = PV(r + 1, i);
    pv_lon[r * p + i] = pv_edge(scheme, order, wgt, v_lon[r * p + i], u_f[r * p + i], 0, dt, M(GMC_LE_LON, r),
                                f1, f2, f3, f4);
  }
  if (r <= g.nhalf - 1) {
    const double f1 = pv[r * p + wrap(i - 2, g.nlon)], f2 = pv[r * p + wrap(i - 1, g.nlon)];
    const double f3 = pv[r * p + i], f4 = pv[r * p + wrap(i + 1, g.nlon)];
    pv_lat[r * p + i] = pv_edge(scheme, order, wgt, u_lat[r * p + i], v_f[r * p + i], 1, dt, M(GMC_LE_LAT, r),
                                f1, f2, f3, f4);
  }
}

// calc_grad_mf, operators_mod.F90:939-1006
__global__ void k_grad_mf(Grid g, Mesh M, const double *__restrict__ mfx, const double *__restrict__ mfy,
                          const double *__restrict__ pole, double *dmfdlon, double *dmfdlat) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch;
  if (r == 0 || r == g.nlat - 1) { dmfdlat[r * p + i] = pole[(r == 0 ? 0 : 1) * 4 + 2]; return; }
  const int im = wrap(i - 1, g.nlon);
  dmfdlon[r * p + i] = dmfdlon_of(mfx[r * p + i], mfx[r * p + im], M, r);
  dmfdlat[r * p + i] = dmfdlat_of(mfy[r * p + i], mfy[(r - 1) * p + i], M, r);
}

// calc_coriolis, operators_mod.F90:846-906
__global__ void k_coriolis(Grid g, Mesh M, int scheme, const double *__restrict__ mfx,
                           const double *__restrict__ mfy, const double *__restrict__ mfx_lat,
                           const double *__restrict__ mfy_lon, const double *__restrict__ pv_lon,
                           const double *__restrict__ pv_lat, double *qhu, double *qhv) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch;
  const int im = wrap(i - 1, g.nlon), ip = wrap(i + 1, g.nlon);
  if (r <= g.nhalf - 1) {
    qhu[r * p + i] = scheme == 1
        ? qhu_of(M, r, pv_lat[r * p + i], mfx[r * p + im], pv_lon[r * p + im], mfx[r * p + i],
                 pv_lon[r * p + i], mfx[(r + 1) * p + im], pv_lon[(r + 1) * p + im],
                 mfx[(r + 1) * p + i], pv_lon[(r + 1) * p + i])
        : mfx_lat[r * p + i] * pv_lat[r * p + i];
  }
  if (r >= 1 && r <= g.nlat - 2) {
    qhv[r * p + i] = scheme == 1
        ? qhv_of(M, r, pv_lon[r * p + i], mfy[(r - 1) * p + i], pv_lat[(r - 1) * p + i],
                 mfy[(r - 1) * p + ip], pv_lat[(r - 1) * p + ip], mfy[r * p + i], pv_lat[r * p + i],
                 mfy[r * p + ip], pv_lat[r * p + ip])
        : mfy_lon[r * p + i] * pv_lon[r * p + i];
  }
}

// calc_grad_ke (:908-937) and the forward-pass combination
// (gmcore_mod.F90:531-548): du = qhv - dkedlon, dv = -qhu - dkedlat,
// dgz = -(dmfdlon + dmfdlat) g.  with_pgf selects the all-pass form.
struct TendPtrs {
  double *du, *dv, *dgz, *qhu, *qhv, *dkedlon, *dkedlat, *dmfdlon, *dmfdlat, *pgf_lon, *pgf_lat;
};
__global__ void k_forward_tend(Grid g, Mesh M, double grav, int with_pgf, const double *__restrict__ ke,
                               TendPtrs T) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, o = r * p + i;
  const int ip = wrap(i + 1, g.nlon);
  if (r >= 1 && r <= g.nlat - 2) {
    const double dk = (ke[r * p + ip] - ke[o]) / M(GMC_DE_LON, r);
    T.dkedlon[o] = dk;
    T.du[o] = with_pgf ? T.qhv[o] - T.pgf_lon[o] - dk : T.qhv[o] - dk;
  }
  if (r <= g.nhalf - 1) {
    const double dk = (ke[o + p] - ke[o]) / M(GMC_DE_LAT, r);
    T.dkedlat[o] = dk;
    T.dv[o] = with_pgf ? -T.qhu[o] - T.pgf_lat[o] - dk : -T.qhu[o] - dk;
  }
  T.dgz[o] = -(T.dmfdlon[o] + T.dmfdlat[o]) * grav;
}

// pgf_swm_run, pgf_swm_mod.F90:18-46
__global__ void k_pgf(Grid g, Mesh M, const double *__restrict__ gz_f, double *pgf_lon, double *pgf_lat) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, o = r * p + i;
  const int ip = wrap(i + 1, g.nlon);
  if (r >= 1 && r <= g.nlat - 2) pgf_lon[o] = pgf_lon_of(gz_f[r * p + ip], gz_f[o], M, r);
  if (r <= g.nhalf - 1) pgf_lat[o] = pgf_lat_of(gz_f[o + p], gz_f[o], M, r);
}

// backward-pass combination, gmcore_mod.F90:557-568
__global__ void k_backward_tend(Grid g, const double *__restrict__ du2, const double *__restrict__ dv2,
                                const double *__restrict__ pgf_lon, const double *__restrict__ pgf_lat,
                                double *du, double *dv) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t o = r * g.pitch + i;
  if (r >= 1 && r <= g.nlat - 2) du[o] = du2[o] - pgf_lon[o];
  if (r <= g.nhalf - 1) dv[o] = dv2[o] - pgf_lat[o];
}

// update_state, gz part (time_schemes_mod.F90:178-183)
__global__ void k_update_gz(Grid g, double dt, const double *__restrict__ gz_old,
                            const double *__restrict__ dgz, double *gz_new) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t o = r * g.pitch + i;
  gz_new[o] = gz_old[o] + dt * dgz[o];
}

// update_state, u/v part with the pole-row v smoothing
// (time_schemes_mod.F90:198-233).  The smoothed rows recompute their
// neighbour's update, which is the same expression.
__global__ void k_update_uv(Grid g, double dt, const double *__restrict__ u_old,
                            const double *__restrict__ v_old, const double *__restrict__ du,
                            const double *__restrict__ dv, double *u_new, double *v_new) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, o = r * p + i;
  if (r >= 1 && r <= g.nlat - 2) u_new[o] = u_old[o] + dt * du[o];
  if (r <= g.nhalf - 1) {
    double vv = v_old[o] + dt * dv[o];
    if (r == 0 || r == g.nhalf - 1) {
      const int64_t q = r == 0 ? o + p : o - p;
      vv = pole_smooth(vv, v_old[q] + dt * dv[q]);
    }
    v_new[o] = vv;
  }
}

// zonal filter (filter_mod.F90:25-163), one workgroup per latitude row.  The
// row is staged in LDS with its periodic halo; each lane forms its taps in
// ascending order.  x == y filters in place (the whole row is read first).
__global__ void __launch_bounds__(256) k_filter(Grid g, int rows, const int32_t *__restrict__ ngrid,
                                                const double *__restrict__ wgt, int max_taps,
                                                const double *x, double *y) {
  extern __shared__ double row[];
  const int r = blockIdx.x;
  if (r >= rows) return;
  const int n = ngrid[r];
  const int64_t o = r * g.pitch;
  if (n <= 1) {
    if (x != y)
      for (int i = threadIdx.x; i < g.nlon; i += blockDim.x) y[o + i] = x[o + i];
    return;
  }
  const int hn = (n - 1) / 2;
  const int len = g.nlon + 2 * hn;
  for (int t = threadIdx.x; t < len; t += blockDim.x) {
    int s = (t - hn) % g.nlon;
    if (s < 0) s += g.nlon;
    row[t] = x[o + s];
  }
  __syncthreads();
  const double *w = wgt + (int64_t)r * max_taps;
  for (int i = threadIdx.x; i < g.nlon; i += blockDim.x) y[o + i] = filter_point(w, n, row + i);
}

// diagnose partial sums (gmcore_mod.F90:230-313): one row per block, the
// host adds the row sums in row order.
__global__ void k_diag_rows(Grid g, Mesh M, double grav, const double *__restrict__ m,
                            const double *__restrict__ gzs, const double *__restrict__ mfx,
                            const double *__restrict__ u, const double *__restrict__ mfy,
                            const double *__restrict__ v, const double *__restrict__ pv,
                            const double *__restrict__ m_lon, const double *__restrict__ pv_lon,
                            const double *__restrict__ m_lat, const double *__restrict__ pv_lat,
                            double *out) {
  const int r = blockIdx.x, lane = threadIdx.x;
  const int64_t o = r * g.pitch;
  const double ac = M(GMC_AREA_CELL, r);
  double s[7];
  s[0] = wave_zonal_sum(g.nlon, lane, [&](int i) { return m[o + i] * ac; });
  s[1] = (r >= 1 && r <= g.nlat - 2) ? wave_zonal_sum(g.nlon, lane, [&](int i) {
    return mfx[o + i] * 0.5 * u[o + i] * M(GMC_AREA_LON, r) * 2; }) : 0.0;
  s[2] = (r <= g.nhalf - 1) ? wave_zonal_sum(g.nlon, lane, [&](int i) {
    return mfy[o + i] * 0.5 * v[o + i] * M(GMC_AREA_LAT, r) * 2; }) : 0.0;
  s[3] = wave_zonal_sum(g.nlon, lane, [&](int i) {
    const double mm = m[o + i];
    return (mm * mm * grav * 0.5 + mm * gzs[o + i]) * ac; });
  s[4] = (r <= g.nhalf - 1) ? wave_zonal_sum(g.nlon, lane, [&](int i) { return pv[o + i] * M(GMC_AREA_VTX, r); }) : 0.0;
  s[5] = (r >= 1 && r <= g.nlat - 2) ? wave_zonal_sum(g.nlon, lane, [&](int i) {
    const double q = pv_lon[o + i]; return m_lon[o + i] * (q * q) * 0.5 * M(GMC_AREA_LON, r); }) : 0.0;
  s[6] = (r <= g.nhalf - 1) ? wave_zonal_sum(g.nlon, lane, [&](int i) {
    const double q = pv_lat[o + i]; return m_lat[o + i] * (q * q) * 0.5 * M(GMC_AREA_LAT, r); }) : 0.0;
  if (lane == 0)
    for (int k = 0; k < 7; k++) out[r * 7 + k] = s[k];
}

}  // namespace

// ---------------------------------------------------------------------------
// host launchers
static Mesh mesh_of(gmc_ctx *c) { return Mesh{c->d_mesh, c->grid.nlat}; }

static void filter_field(gmc_ctx *c, const FilterDev &f, int half, const double *x, double *y) {
  const Grid &g = c->grid;
  const int rows = half ? g.nhalf : g.nlat;
  const size_t lds = sizeof(double) * (g.nlon + f.max_taps + 1);
  hipLaunchKernelGGL(k_filter, dim3(rows), dim3(256), lds, c->stream, g, rows,
                     half ? f.ngrid_lat : f.ngrid_lon, half ? f.wgt_lat : f.wgt_lon, f.max_taps, x, y);
}

static TendPtrs tend_ptrs(double **T) {
  return TendPtrs{T[T_DU], T[T_DV], T[T_DGZ], T[T_QHU], T[T_QHV], T[T_DKEDLON], T[T_DKEDLAT],
                  T[T_DMFDLON], T[T_DMFDLAT], T[T_PGF_LON], T[T_PGF_LAT]};
}

// operators_prepare_2's forward pass on state S with the stage dt; pole:
// 8 doubles for the pole sums (k_pole_sums)
static void forward_prepare(gmc_ctx *c, double **S, double dt, double *pole) {
  const Grid &g = c->grid;
  const Mesh M = mesh_of(c);
  const dim3 b(BX, BY);
  hipLaunchKernelGGL(k_calc_mf, grid2d(g, g.nlat), b, 0, c->stream, g, S[S_M_LON], S[S_M_LAT], S[S_U_F],
                     S[S_V_F], M, S[S_MFX_LON], S[S_MFY_LAT], S[S_MFX_LAT], S[S_MFY_LON], S[S_U_LAT],
                     S[S_V_LON]);
  hipLaunchKernelGGL(k_pole_sums, dim3(2), dim3(kWave), 0, c->stream, g, M, S[S_V_F], S[S_MFY_LAT],
                     S[S_U_LAT], pole);
  hipLaunchKernelGGL(k_calc_ke, grid2d(g, g.nlat), b, 0, c->stream, g, M, S[S_U_F], S[S_V_F], pole,
                     S[S_KE]);
  hipLaunchKernelGGL(k_calc_div, grid2d(g, g.nlat), b, 0, c->stream, g, M, S[S_U_F], S[S_V_F], pole,
                     S[S_DIV]);
  hipLaunchKernelGGL(k_calc_pv, grid2d(g, g.nhalf), b, 0, c->stream, g, M, c->cfg.pv_pole_stokes,
                     S[S_U_F], S[S_V_F], S[S_M_VTX], pole, S[S_VOR], S[S_PV]);
  const int order = c->cfg.pv_scheme == GMC_PV_TVD ? c->cfg.tvd_limiter : c->cfg.upwind_order_pv;
  hipLaunchKernelGGL(k_interp_pv, grid2d(g, g.nlat), b, 0, c->stream, g, M, c->cfg.pv_scheme, order,
                     c->cfg.upwind_wgt_pv, dt, S[S_PV], S[S_U_F], S[S_V_F], S[S_U_LAT], S[S_V_LON],
                     S[S_PV_LON], S[S_PV_LAT]);
}

static void calc_m(gmc_ctx *c, double **S) {
  const Grid &g = c->grid;
  hipLaunchKernelGGL(k_calc_m, grid2d(g, g.nlat), dim3(BX, BY), 0, c->stream, g, mesh_of(c), c->cfg.g,
                     S[S_GZ], c->gzs, S[S_M], S[S_M_LON], S[S_M_LAT], S[S_M_VTX]);
}

int ops_prepare(gmc_ctx *c, int level) {
  double **S = c->state[level];
  filter_field(c, c->big, 0, S[S_U_LON], S[S_U_F]);
  filter_field(c, c->big, 1, S[S_V_LAT], S[S_V_F]);
  filter_field(c, c->big, 0, S[S_GZ], S[S_GZ_F]);
  calc_m(c, S);
  forward_prepare(c, S, c->cfg.dt_dyn, c->d_pole);   // operators_prepare_1 with dt_dyn (gmcore_mod.F90:133)
  return check_hip(c, hipGetLastError(), "prepare");
}

int ops_space_operators(gmc_ctx *c, int old_l, int star_l, int new_l, int t1, int t2, double dt, int pass) {
  (void)old_l;
  const Grid &g = c->grid;
  const Mesh M = mesh_of(c);
  const dim3 b(BX, BY);
  double **S = c->state[star_l], **N = c->state[new_l], **T1 = c->tend[t1], **T2 = c->tend[t2];
  if (pass == 0 || pass == 1) {
    if (pass == 0) calc_m(c, S);
    forward_prepare(c, S, dt, c->d_pole);
    hipLaunchKernelGGL(k_grad_mf, grid2d(g, g.nlat), b, 0, c->stream, g, M, S[S_MFX_LON], S[S_MFY_LAT],
                       c->d_pole, T1[T_DMFDLON], T1[T_DMFDLAT]);
    hipLaunchKernelGGL(k_coriolis, grid2d(g, g.nlat), b, 0, c->stream, g, M, c->cfg.coriolis_scheme,
                       S[S_MFX_LON], S[S_MFY_LAT], S[S_MFX_LAT], S[S_MFY_LON], S[S_PV_LON], S[S_PV_LAT],
                       T1[T_QHU], T1[T_QHV]);
    if (pass == 0)
      hipLaunchKernelGGL(k_pgf, grid2d(g, g.nlat), b, 0, c->stream, g, M, S[S_GZ_F], T1[T_PGF_LON],
                         T1[T_PGF_LAT]);
    hipLaunchKernelGGL(k_forward_tend, grid2d(g, g.nlat), b, 0, c->stream, g, M, c->cfg.g, pass == 0 ? 1 : 0,
                       S[S_KE], tend_ptrs(T1));
  } else if (pass == 2) {
    calc_m(c, N);
    hipLaunchKernelGGL(k_pgf, grid2d(g, g.nlat), b, 0, c->stream, g, M, N[S_GZ_F], T1[T_PGF_LON],
                       T1[T_PGF_LAT]);
    hipLaunchKernelGGL(k_backward_tend, grid2d(g, g.nlat), b, 0, c->stream, g, T2[T_DU], T2[T_DV],
                       T1[T_PGF_LON], T1[T_PGF_LAT], T1[T_DU], T1[T_DV]);
  } else {
    return set_error(c, GMC_EINVAL, "pass must be 0, 1 or 2");
  }
  return check_hip(c, hipGetLastError(), "space_operators");
}

int ops_update_state(gmc_ctx *c, int tend, int old_l, int new_l, double dt, int what) {
  const Grid &g = c->grid;
  const dim3 b(BX, BY);
  double **O = c->state[old_l], **N = c->state[new_l], **T = c->tend[tend];
  if (what & 1) {
    hipLaunchKernelGGL(k_update_gz, grid2d(g, g.nlat), b, 0, c->stream, g, dt, O[S_GZ], T[T_DGZ], N[S_GZ]);
    filter_field(c, c->big, 0, N[S_GZ], N[S_GZ_F]);
    filter_field(c, c->small1, 0, N[S_GZ], N[S_GZ]);
  }
  if (what & 2) {
    hipLaunchKernelGGL(k_update_uv, grid2d(g, g.nlat), b, 0, c->stream, g, dt, O[S_U_LON], O[S_V_LAT],
                       T[T_DU], T[T_DV], N[S_U_LON], N[S_V_LAT]);
    filter_field(c, c->big, 0, N[S_U_LON], N[S_U_F]);
    filter_field(c, c->big, 1, N[S_V_LAT], N[S_V_F]);
  }
  return check_hip(c, hipGetLastError(), "update_state");
}

// Zonal filter of one field through the context's stream (x == y allowed).
void ops_filter(gmc_ctx *c, int small, int half, const double *x, double *y) {
  filter_field(c, small ? c->small1 : c->big, half, x, y);
}

// One step_forward_backward stage old -> new with star `s` on the operator
// kernels, leaving every intermediate in place: m and the operators_prepare
// fields of the star state, the forward tendencies in tend level 1, the
// backward ones in tend level 2, the new state in `n`.  m of the star state
// is recomputed first (the same values gmc_prepare / the previous backward
// pass left there).
int ops_stage_tape(gmc_ctx *c, int o, int s, int n, double dt) {
  calc_m(c, c->state[s]);
  int rc = ops_space_operators(c, o, s, n, 1, 2, dt, 1);
  if (!rc) rc = ops_update_state(c, 1, o, n, dt, 1);
  if (!rc) rc = ops_space_operators(c, o, s, n, 2, 1, dt, 2);
  if (!rc) rc = ops_update_state(c, 2, o, n, dt, 2);
  return rc;
}

// div_damp_run order 2 (div_damp_mod.F90:89-139)
namespace {
// div2 of calc_div for div_damp_order 4 (operators_mod.F90:414-428): the
// Laplacian of div on the non-pole rows; the pole rows of div2 stay zero.
__global__ void k_calc_div2(Grid g, Mesh M, const double *__restrict__ div, double *div2) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r < 1 || r > g.nlat - 2) return;
  const int64_t p = g.pitch, o = r * p + i;
  const int ip = wrap(i + 1, g.nlon), im = wrap(i - 1, g.nlon);
  const double dl = M(GMC_DE_LON, r);
  div2[o] = (div[r * p + ip] - 2 * div[o] + div[r * p + im]) / (dl * dl) +
            ((div[o + p] - div[o]) * M(GMC_HALF_COS_LAT, r) / M(GMC_DE_LAT, r) -
             (div[o] - div[o - p]) * M(GMC_HALF_COS_LAT, r - 1) / M(GMC_DE_LAT, r - 1)) /
                M(GMC_LE_LON, r) / M(GMC_FULL_COS_LAT, r);
}

// div_damp_run (div_damp_mod.F90:89-139) with the c_lon / c_lat of
// div_damp_init (:27-80): order 2 adds the gradient of div, order 4
// subtracts the gradient of div2 (d is div or div2).
__global__ void k_div_damp(Grid g, Mesh M, int order, double coef, const double *__restrict__ d, double *u,
                           double *v) {
  const int i = blockIdx.x * BX + threadIdx.x, r = blockIdx.y * BY + threadIdx.y;
  if (i >= g.nlon || r >= g.nlat) return;
  const int64_t p = g.pitch, o = r * p + i;
  const int ip = wrap(i + 1, g.nlon);
  if (r >= 1 && r <= g.nlat - 2) {
    const double le = M(GMC_LE_LON, r), de = M(GMC_DE_LON, r);
    const double c = order == 4 ? coef * (le * le) * (de * de) : coef * le * de;
    const double t = c * (d[r * p + ip] - d[o]) / de;
    u[o] = order == 4 ? u[o] - t : u[o] + t;
  }
  if (r <= g.nhalf - 1) {
    const double le = M(GMC_LE_LAT, r), de = M(GMC_DE_LAT, r);
    const double c = order == 4 ? coef * (le * le) * (de * de) : coef * le * de;
    const double t = c * (d[o + p] - d[o]) / de;
    v[o] = order == 4 ? v[o] - t : v[o] + t;
  }
}
}  // namespace

int ops_calc_div(gmc_ctx *c, int level) {
  const Grid &g = c->grid;
  const Mesh M = mesh_of(c);
  double **S = c->state[level];
  hipLaunchKernelGGL(k_pole_sums, dim3(2), dim3(kWave), 0, c->stream, g, M, S[S_V_F], S[S_MFY_LAT],
                     S[S_U_LAT], c->d_pole);
  hipLaunchKernelGGL(k_calc_div, grid2d(g, g.nlat), dim3(BX, BY), 0, c->stream, g, M, S[S_U_F], S[S_V_F],
                     c->d_pole, S[S_DIV]);
  return check_hip(c, hipGetLastError(), "calc_div");
}

int ops_div_damp(gmc_ctx *c, int level) {
  const Grid &g = c->grid;
  const Mesh M = mesh_of(c);
  const dim3 b(BX, BY);
  double **S = c->state[level];
  // calc_div needs the pole sums of the filtered v of this state
  hipLaunchKernelGGL(k_pole_sums, dim3(2), dim3(kWave), 0, c->stream, g, M, S[S_V_F], S[S_MFY_LAT],
                     S[S_U_LAT], c->d_pole);
  hipLaunchKernelGGL(k_calc_div, grid2d(g, g.nlat), b, 0, c->stream, g, M, S[S_U_F], S[S_V_F], c->d_pole,
                     S[S_DIV]);
  const int o4 = c->cfg.div_damp_order == 4;
  if (o4) hipLaunchKernelGGL(k_calc_div2, grid2d(g, g.nlat), b, 0, c->stream, g, M, S[S_DIV], S[S_DIV2]);
  hipLaunchKernelGGL(k_div_damp, grid2d(g, g.nlat), b, 0, c->stream, g, M, o4 ? 4 : 2,
                     o4 ? c->cfg.div_damp_coef4 : c->cfg.div_damp_coef2, o4 ? S[S_DIV2] : S[S_DIV],
                     S[S_U_LON], S[S_V_LAT]);
  return check_hip(c, hipGetLastError(), "div_damp");
}

// The launches diagnose shares with gmc_diagnose_dev: m and, where the
// level does not hold them, the forward-pass fields (pole sums into pole[0..7]),
// then one row of partial sums per latitude row into part[r * 7 + k].
static int diag_rows(gmc_ctx *c, int level, double *pole, double *part) {
  const Grid &g = c->grid;
  double **S = c->state[level];
  // m, m_lon, m_lat as the backward pass's calc_m left them on the current
  // gz; the forward-pass fields as the level's last use as a star state left
  // them (the reference reads them as stored, gmcore_mod.F90:230-313), or,
  // when the level does not hold them, formed now from its own state
  calc_m(c, S);
  if (!c->derived_kept[level]) forward_prepare(c, S, c->cfg.dt_dyn, pole);
  hipLaunchKernelGGL(k_diag_rows, dim3(g.nlat), dim3(kWave), 0, c->stream, g, mesh_of(c), c->cfg.g, S[S_M],
                     c->gzs, S[S_MFX_LON], S[S_U_LON], S[S_MFY_LAT], S[S_V_LAT], S[S_PV], S[S_M_LON],
                     S[S_PV_LON], S[S_M_LAT], S[S_PV_LAT], part);
  return check_hip(c, hipGetLastError(), "diagnose");
}

int ops_diagnose(gmc_ctx *c, int level, double out[4]) {
  const Grid &g = c->grid;
  int rc = diag_rows(c, level, c->d_pole, c->d_diag);
  if (rc) return rc;
  double *h = new double[(size_t)g.nlat * 7];
  rc = check_hip(c, hipMemcpyAsync(h, c->d_diag, sizeof(double) * g.nlat * 7, hipMemcpyDeviceToHost, c->stream),
                 "diagnose copy");
  if (!rc) rc = check_hip(c, hipStreamSynchronize(c->stream), "diagnose sync");
  if (!rc) {
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int r = 0; r < g.nlat; r++)
      for (int k = 0; k < 7; k++) s[k] += h[r * 7 + k];
    out[0] = s[0];
    out[1] = s[1] + s[2] + s[3];
    out[2] = s[5] + s[6];
    out[3] = s[4];
  }
  delete[] h;
  return rc;
}

// ---------------------------------------------------------------------------
// gmc_diagnose_dev: diagnose without allocating.  Its pole sums and row
// partial sums live in `scratch`, a static device array of this file (one
// copy per GPU, kept until the process ends); the row sums are added on the
// device in the host loop's order, so out[] holds gmc_diagnose's bits.
// Layout: [0, 64) pole sums (k_pole_sums writes 8 of them), [64, 64 + 7 nlat)
// row partials.  Every element a call reads is written earlier in the same
// call, so nothing needs clearing between calls.
constexpr int kScratchMaxRows = 4097;   // latitude rows, poles included
constexpr size_t kScratchDoubles = 64 + 7 * (size_t)kScratchMaxRows;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

namespace {
// the totals of diagnose from the row partials, in the order of ops_diagnose's
// host loop (row r ascending for each sum, then the same combinations)
__global__ void k_diag_total(int nlat, const double *__restrict__ part, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int r = 0; r < nlat; r++)
    for (int k = 0; k < 7; k++) s[k] += part[r * 7 + k];
  out[0] = s[0];
  out[1] = s[1] + s[2] + s[3];
  out[2] = s[5] + s[6];
  out[3] = s[4];
}
}  // namespace

int ops_scratch_init(gmc_ctx *c) {
  int dev = 0;
  int rc = check_hip(c, hipGetDevice(&dev), "get device");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices) return set_error(c, GMC_EDEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    rc = check_hip(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch symbol address");
    if (rc) return rc;
    g_scratch[dev] = (double *)ptr;
  }
  c->scratch = g_scratch[dev];
  return GMC_OK;
}

int ops_diagnose_dev(gmc_ctx *c, int level, double *d_out) {
  const Grid &g = c->grid;
  if (g.nlat > kScratchMaxRows)
    return set_error(c, GMC_EINVAL, "gmc_diagnose_dev: more latitude rows than its static scratch holds (4097)");
  if (!c->scratch) return set_error(c, GMC_ESTATE, "gmc_diagnose_dev: scratch address not resolved");
  double *pole = c->scratch, *part = c->scratch + 64;
  int rc = diag_rows(c, level, pole, part);
  if (rc) return rc;
  hipLaunchKernelGGL(k_diag_total, dim3(1), dim3(kWave), 0, c->stream, g.nlat, part, d_out);
  return check_hip(c, hipGetLastError(), "diagnose total");
}

}  // namespace gmc
