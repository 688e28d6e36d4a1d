"""Pins of the zonal polar filter tables (filter_types_mod.F90:31-147).

- tap counts and totals equal the hand-computed table of the configurations
  (widths from the CFL formula with the namelist defaults, including the
  single-precision literals 0.4 and 0.2 of namelist_mod.F90:92-93);
- weights are the closed-form Gaussian exp(-x^2 / (2 s^2)), s = b w / 2,
  normalised: sum 1, symmetric, ratio to the centre tap exact to 1e-15,
  with the widths of small_filter1 (x filter_coef_d) and small_filter2 (x
  exp_two_values of filter_coef_e and filter_coef_d) and their tap counts;
- the Fortran host and the C oracle build identical tables;
- filtering a zonally constant field returns the constant (through the
  oracle's filter application at gmcore_run's set-up).
"""
import numpy as np
import pytest

import oracle
from gmcore_amd.capi import MESH
from gmcore_amd.host import Filter, Mesh, Setup
from gmcore_amd.namelist import Namelist

# (nlon, nlat, dt): big n_max full / half rows, sum of taps, rows with small n > 1, halo
TABLE = [
    ((180, 91, 600.0), (49, 91, 1523, 32, 45)),
    ((360, 181, 300.0), (91, 169, 3267, 64, 84)),
    ((720, 361, 150.0), (169, 317, 6971, 130, 158)),
    ((1440, 721, 60.0), (255, 493, 12991, 204, 246)),
]


@pytest.mark.parametrize("grid,expect", TABLE, ids=lambda x: str(x))
def test_tap_table(grid, expect):
    nlon, nlat, dt = grid
    nl = Namelist()
    m = oracle.mesh_metrics(nlon, nlat, nl.radius, nl.omega)
    big = oracle.filter_table(nlon, nlat, m, dt, nl, 0)
    small = oracle.filter_table(nlon, nlat, m, dt, nl, 1)
    assert (big[0].max(), big[2].max()) == expect[:2]
    assert int(big[0].sum() + big[2].sum()) == expect[2]
    assert int((small[0] > 1).sum()) == expect[3]
    assert max(2, int((big[2].max() - 1) // 2)) == expect[4]
    assert np.all((big[0][1:-1] % 2) == 1) and big[0][0] == 0 and big[0][-1] == 0


def test_namelist_single_precision_defaults():
    nl = Namelist()
    assert nl.filter_coef_b == float(np.float32(0.4)) != 0.4
    assert nl.filter_coef_c == float(np.float32(0.2)) != 0.2
    assert nl.filter_coef_d == 0.375 and nl.max_cfl == 0.5


def small2_factor(nl, lat, x0):
    """exp_two_values(filter_coef_e, filter_coef_d, x0, 80, |lat|) (math_mod.F90:146-159)"""
    w = np.exp((abs(lat) - x0) ** 2 * np.log(1.0e-3) / (80.0 - x0) ** 2)
    return w * nl.filter_coef_e + (1 - w) * nl.filter_coef_d


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_weights_closed_form(kind):
    nlon, nlat, dt = 360, 181, 300.0
    nl = Namelist()
    m = oracle.mesh_metrics(nlon, nlat, nl.radius, nl.omega)
    ng_lon, w_lon, ng_lat, w_lat = oracle.filter_table(nlon, nlat, m, dt, nl, kind)
    big = oracle.filter_table(nlon, nlat, m, dt, nl, 0)
    checked = 0
    for ng, w, dx, lat, x0, nb in ((ng_lon, w_lon, m[MESH["de_lon"]], m[MESH["full_lat_deg"]],
                                    -m[MESH["full_lat_deg"], 1], big[0]),
                                   (ng_lat, w_lat, m[MESH["le_lat"], :nlat - 1], m[MESH["half_lat_deg"], :nlat - 1],
                                    -m[MESH["half_lat_deg"], 0], big[2])):
        for r in range(len(ng)):
            n = int(ng[r])
            if n <= 1:
                continue
            ww = w[r, :n]
            assert abs(ww.sum() - 1) <= 4e-16 * n
            assert np.array_equal(ww, ww[::-1])
            # width of the row from the CFL formula
            width = nl.filter_coef_a * (nl.max_wave_speed * dt / dx[r]) / nl.max_cfl * (
                nl.filter_coef_c * (np.tanh(90 - abs(lat[r])) - 1) + 1)
            if kind == 1:
                width *= nl.filter_coef_d
            if kind == 2:
                f = small2_factor(nl, lat[r], x0)
                width *= f
                want = int(f * nb[r])
                assert n == (want + 1 if want % 2 == 0 else want)
            s = nl.filter_coef_b * width / 2
            x = np.arange(n) - (n - 1) // 2
            assert np.allclose(ww / ww[(n - 1) // 2], np.exp(-x ** 2 / (2 * s * s)), rtol=1e-13, atol=0)
            checked += 1
    assert checked > 50


def test_small_filter2_factor_ends():
    """small_filter2 shrinks the big filter by filter_coef_e next to the poles
    and by about filter_coef_d from 80 degrees to the equator."""
    nl = Namelist()
    assert small2_factor(nl, 89.0, 89.0) == nl.filter_coef_e
    assert small2_factor(nl, 80.0, 89.0) == pytest.approx(1e-3 * nl.filter_coef_e + 0.999 * nl.filter_coef_d,
                                                          rel=1e-15)
    nlon, nlat, dt = 360, 181, 300.0
    m = oracle.mesh_metrics(nlon, nlat, nl.radius, nl.omega)
    big = oracle.filter_table(nlon, nlat, m, dt, nl, 0)
    s2 = oracle.filter_table(nlon, nlat, m, dt, nl, 2)
    rows = np.nonzero(big[0] > 1)[0]
    assert np.all(s2[0][rows] <= big[0][rows]) and np.all(s2[0][rows] % 2 == 1)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_fortran_host_filter_bitwise(kind):
    nlon, nlat, dt = 180, 91, 600.0
    nl = Namelist()
    m = oracle.mesh_metrics(nlon, nlat, nl.radius, nl.omega)
    o = oracle.filter_table(nlon, nlat, m, dt, nl, kind)
    h = Filter(Mesh(nlon, nlat, nl.radius, nl.omega), dt, nl, kind)
    assert np.array_equal(o[0], h.ngrid_lon) and np.array_equal(o[2], h.ngrid_lat)
    n = h.max_taps
    assert np.array_equal(o[1][:, :n], h.wgt_lon) and np.array_equal(o[3][:, :n], h.wgt_lat)


def test_filter_of_constant_is_constant():
    nl = Namelist(num_lon=180, num_lat=91, dt_dyn=600.0, time_scheme="pc2")
    st = Setup(nl)
    o = oracle.SWOracle(st.mesh_struct, st.cfg)
    rng = np.random.default_rng(3)
    c_u = rng.uniform(-5, 5, size=(91, 1)) * np.ones((1, 180))
    c_v = rng.uniform(-5, 5, size=(90, 1)) * np.ones((1, 180))
    o.set(1, "u_lon", c_u)
    o.set(1, "v_lat", c_v)
    o.set(1, "gz", np.full((91, 180), 5e4))
    o.set(1, "gzs", np.zeros((91, 180)))
    o.prepare(1)
    assert np.allclose(o.get(1, "u_f"), c_u, rtol=1e-15, atol=0)
    assert np.allclose(o.get(1, "v_f"), c_v, rtol=1e-15, atol=0)
    assert np.allclose(o.get(1, "gz_f"), 5e4, rtol=1e-15, atol=0)
