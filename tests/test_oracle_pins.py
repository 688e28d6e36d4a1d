"""Pins of oracle parts that no reference fixture covers: divergence damping
(orders 2 and 4), the diagnose sums and the tvd edge values, each against
what the mathematics fixes.

- Divergence damping (div_damp_mod.F90:27-139, calc_div operators_mod.F90:
  354-431): a discretely nondivergent wind is left unchanged; order 2 never
  raises the area-weighted divergence variance (summation by parts makes
  Σ div·Δdiv·A = -Σ c·(grad div)² ≤ 0); a single zonal wavenumber in one
  row decays by the closed-form factor of its difference operators.  A
  flipped sign, a wrong metric or a shifted index fails one of them.
- diagnose (gmcore_mod.F90:213-341): at rest tm = Φ0/g·ΣA and
  te = Φ0²/(2g)·ΣA exactly, tav vanishes by the equatorial antisymmetry of
  f; Σ vor·A_vtx telescopes to zero (discrete Stokes) without the pole
  circulation override.
- tvd (tvd_mod.F90:42-100): with limiter 'none' the Lax-Wendroff value;
  at an extremum van Leer and MC fall back to first-order upwind; on linear
  data (r = 1) both give Lax-Wendroff; MC saturates at 2 and follows
  (1 + r)/2 in between (van Leer 1977's monotonised-central limiter).
"""
import math

import numpy as np
import pytest

import oracle
from helpers import make_oracle, namelist, seeded_ic
from gmcore_amd.host import Setup

NLON, NLAT = 72, 37


def _setup(**kw):
    return Setup(namelist(NLON, NLAT, 1800.0, **kw))


def _with_winds(st, u, v, gz=None):
    """oracle with u_lon = u_f = u, v_lat = v_f = v on level 1"""
    gz = np.full((NLAT, NLON), 5.0e4) if gz is None else gz
    o = make_oracle(st, (u, v, gz, np.zeros((NLAT, NLON))))
    o.set(1, "u_f", u)
    o.set(1, "v_f", v)
    return o


@pytest.mark.parametrize("order", [2, 4])
def test_div_damp_leaves_a_nondivergent_wind(order):
    st = _setup(use_div_damp=True, div_damp_order=order)
    u, v, _, _ = seeded_ic(st, lmax=8, seed=3)   # u, v from a vertex streamfunction
    o = _with_winds(st, u, v)
    o.div_damp(1)
    du = np.abs(o.get(1, "u_lon") - u)[1:-1].max()
    dv = np.abs(o.get(1, "v_lat") - v).max()
    scale = max(np.abs(u).max(), np.abs(v).max())
    assert du <= 1e-13 * scale and dv <= 1e-13 * scale, (du, dv)
    # a divergent wind of the same size does change
    rng = np.random.default_rng(1)
    ur = rng.standard_normal(u.shape) * scale
    ur[[0, -1]] = 0.0
    o2 = _with_winds(st, ur, rng.standard_normal(v.shape) * scale)
    o2.div_damp(1)
    assert np.abs(o2.get(1, "u_lon") - ur)[1:-1].max() > 1e-6 * scale


def test_div_damp_order2_never_raises_divergence_variance():
    st = _setup(use_div_damp=True, div_damp_order=2)
    rng = np.random.default_rng(2)
    u = rng.standard_normal((NLAT, NLON))
    u[[0, -1]] = 0.0
    v = rng.standard_normal((NLAT - 1, NLON))
    o = _with_winds(st, u, v)
    o.div_damp(1)
    div0 = o.internal(1, "div")
    o.set(1, "u_f", o.get(1, "u_lon"))
    o.set(1, "v_f", o.get(1, "v_lat"))
    o.div_damp(1)
    div1 = o.internal(1, "div")
    area = st.mesh["area_cell"][:, None]
    s = float(np.sum(div0 * (div1 - div0) * area))
    var = float(np.sum(div0 * div0 * area))
    assert s < -1e-4 * var, (s, var)


@pytest.mark.parametrize("order,k", [(2, 1), (2, 7), (4, 3), (4, 11)])
def test_div_damp_zonal_mode_decays_by_its_closed_form(order, k):
    coef2, coef4 = 1.0 / 128.0, 0.01
    st = _setup(use_div_damp=True, div_damp_order=order)
    j0 = 12                                   # a mid-latitude full row (0-based)
    dl = st.mesh.dlon
    lam_u = (np.arange(NLON) + 0.5) * dl      # u points sit half a cell east
    u = np.zeros((NLAT, NLON))
    u[j0] = np.cos(k * lam_u)
    v = np.zeros((NLAT - 1, NLON))
    o = _with_winds(st, u, v)
    o.div_damp(1)
    M = st.mesh
    le, A, de = M["le_lon"][j0], M["area_cell"][j0], M["de_lon"][j0]
    s2 = math.sin(k * dl / 2) ** 2
    if order == 2:
        factor = 1 - 4 * coef2 * le * le / A * s2
    else:
        mu = (M["half_cos_lat"][j0] / M["de_lat"][j0] + M["half_cos_lat"][j0 - 1] / M["de_lat"][j0 - 1]) \
            / (le * M["full_cos_lat"][j0])
        factor = 1 - coef4 * le * le * de * (le / A) * (16 * s2 * s2 / (de * de) + 4 * mu * s2)
    got = o.get(1, "u_lon")
    assert abs(factor - 1) > 1e-6
    assert np.abs(got[j0] - factor * u[j0]).max() <= 1e-12, (factor, got[j0][:3] / u[j0][:3])
    near = [j0] if order == 2 else [j0 - 1, j0, j0 + 1]
    if order == 4:
        # div2 of the neighbour rows is the meridional part of the Laplacian
        # of row j0's div alone, so their u gets c * nu * (le/A) * 4 s^2 * u0
        for r, hrow in ((j0 + 1, j0), (j0 - 1, j0 - 1)):
            nu = M["half_cos_lat"][hrow] / M["de_lat"][hrow] / (M["le_lon"][r] * M["full_cos_lat"][r])
            want = coef4 * M["le_lon"][r] ** 2 * M["de_lon"][r] * nu * (le / A) * 4 * s2 * u[j0]
            assert np.abs(got[r] - want).max() <= 1e-12, r
    others = np.delete(got, near, axis=0)[1:-1]
    assert np.abs(others).max() == 0.0


def test_diagnose_at_rest():
    phi0 = 5.0e4
    st = _setup()
    g = st.nl.g
    o = _with_winds(st, np.zeros((NLAT, NLON)), np.zeros((NLAT - 1, NLON)), np.full((NLAT, NLON), phi0))
    o.prepare(1)
    tm, te, tpe, tav = o.diagnose()
    area = NLON * float(np.sum(st.mesh["area_cell"]))
    assert abs(tm - phi0 / g * area) <= 1e-13 * tm
    assert abs(te - phi0 * phi0 / (2 * g) * area) <= 1e-13 * te
    pv = o.internal(1, "pv")
    scale = float(np.sum(np.abs(pv) * st.mesh["area_vtx"][:NLAT - 1, None]))
    assert scale > 0 and abs(tav) <= 1e-13 * scale, (tav, scale)
    assert tpe > 0


def test_vorticity_sums_to_zero_over_the_sphere():
    st = _setup(pv_pole_stokes=False)
    ic = seeded_ic(st, lmax=8, seed=5)
    o = make_oracle(st, ic)
    o.prepare(1)
    vor = o.internal(1, "vor")
    w = vor * st.mesh["area_vtx"][:NLAT - 1, None]
    assert abs(float(np.sum(w))) <= 1e-13 * float(np.sum(np.abs(w)))


def _lw(cfl, f, fp1):
    return 0.5 * (fp1 + f - cfl * (fp1 - f))


@pytest.mark.parametrize("cfl", [0.3, -0.45])
def test_tvd_limiters(cfl):
    fm1, f, fp1, fp2 = 0.7, 1.9, 1.1, 2.6     # extremum at f and at fp1: r < 0
    assert oracle.tvd("none", cfl, fm1, f, fp1, fp2) == pytest.approx(_lw(cfl, f, fp1), abs=1e-15)
    up = f if cfl > 0 else fp1
    for lim in ("van_leer", "mc"):
        assert oracle.tvd(lim, cfl, fm1, f, fp1, fp2) == pytest.approx(up, abs=1e-15), lim
        # linear data: r = 1, limiter 1, Lax-Wendroff
        assert oracle.tvd(lim, cfl, 0.0, 1.0, 2.0, 3.0) == pytest.approx(_lw(cfl, 1.0, 2.0), abs=1e-15), lim
    # MC (cfl > 0, r = (f - fm1) / (fp1 - f) up to eps): r = 10 saturates at
    # C = 2; r = 1.5 gives C = (1 + r) / 2
    if cfl > 0:
        for fm1_, C in ((-9.0, 2.0), (-0.5, 1.25)):
            expect = 0.5 * (2.0 + 1.0 - ((1 - C) * 1.0 + cfl * C) * 1.0)
            assert oracle.tvd("mc", cfl, fm1_, 1.0, 2.0, 2.5) == pytest.approx(expect, rel=1e-5)


def _response(w, n, k, dl):
    """frequency response of a centred symmetric filter of n taps at zonal wavenumber k"""
    h = (n - 1) // 2
    return float(sum(w[m] * math.cos(k * (m - h) * dl) for m in range(n)))


def test_update_state_filter_sequence():
    """update_state (time_schemes_mod.F90:178-239) on a state with zero
    tendencies: gz_f is the big filter of the unfiltered new gz and gz the
    small filter of it (not one filter after the other), u_f and v_f the big
    filters of u and v with their own (U-row, V-row) tables.  On a single
    zonal wavenumber each filter multiplies the mode by its frequency
    response, formed here from the filter tables (pinned separately); a
    swapped order or table gives the product or the other table's
    response instead."""
    st = _setup()
    dl = st.mesh.dlon
    big, small = st.big, st.small
    r = 3                                     # a polar full row with both filters on
    assert big.ngrid_lon[r] > 1 and small.ngrid_lon[r] > 1
    hr = 2                                    # a polar half row with the big filter on
    assert big.ngrid_lat[hr] > 1
    k, a, phi0 = NLON // 6, 100.0, 5.0e4
    lam = np.arange(NLON) * dl
    gz = np.full((NLAT, NLON), phi0)
    gz[r] += a * np.cos(k * lam)
    u = np.zeros((NLAT, NLON))
    u[r] = np.cos(k * (lam + dl / 2))
    v = np.zeros((NLAT - 1, NLON))
    v[hr] = np.sin(k * lam)
    o = _with_winds(st, u, v, gz)
    o.update_state(1, 1, 2, 1800.0, 3)        # tend level 1 is all zero
    B = _response(big.wgt_lon[r], big.ngrid_lon[r], k, dl)
    S = _response(small.wgt_lon[r], small.ngrid_lon[r], k, dl)
    Bu = B
    Bv = _response(big.wgt_lat[hr], big.ngrid_lat[hr], k, dl)
    assert min(abs(B - S), abs(B - B * S), abs(S - B * S), abs(B - Bv)) > 1e-3, (B, S, Bv)
    gz_f, gz_new = o.get(2, "gz_f"), o.get(2, "gz")
    assert np.abs(gz_f[r] - phi0 - B * a * np.cos(k * lam)).max() <= 1e-9
    assert np.abs(gz_new[r] - phi0 - S * a * np.cos(k * lam)).max() <= 1e-9
    assert np.abs(o.get(2, "u_f")[r] - Bu * u[r]).max() <= 1e-12
    assert np.abs(o.get(2, "v_f")[hr] - Bv * v[hr]).max() <= 1e-12
    assert np.array_equal(o.get(2, "u_lon"), u) and np.array_equal(o.get(2, "v_lat")[1:-1], v[1:-1])
