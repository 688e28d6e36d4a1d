"""Pins of the oracle's shallow-water operators and time stepping (CPU).

- rest state: u = v = 0, constant gz -> every tendency exactly zero, state
  unchanged bit for bit after a pc2 step;
- flux form: sum(dgz * area_cell) = 0 to rounding, pole rows included, and
  total mass tm conserved over many steps;
- the energy identities of debug_check_space_operators
  (src/utils/debug_mod.F90:182-222): Coriolis does no work (ip_cf), the
  kinetic-energy (ip_ke) and potential-energy (ip_pe) exchanges cancel; each
  is checked relative to the sum of the magnitudes of its terms, which makes
  a wrong sign, tangent weight or pole formula fail;
- hemispheric mirror symmetry: a flow mirrored about the equator (u even,
  v odd, gz even) stays mirrored;
- steady geostrophic flow (Williamson case 2,
  src/tests/swm/steady_geostrophic_flow_test_mod.F90) stays near steady and
  its error falls with resolution;
- pc2 composition: gmo_sw_step equals the three step_forward_backward calls
  of time_schemes_mod.F90:244-258 driven from here.
"""
import numpy as np
import pytest

import oracle
from gmcore_amd import init_cond
from gmcore_amd.capi import MESH
from gmcore_amd.host import Setup
from gmcore_amd.namelist import Namelist

from helpers import make_oracle, namelist, seeded_ic


def _setup(nlon=72, nlat=37, dt=1800.0, **kw):
    return Setup(namelist(nlon, nlat, dt, **kw))


def test_rest_state_is_exact():
    """With the filters off a resting, flat state has identically zero
    tendencies and does not move."""
    st = _setup(max_wave_speed=0.0)
    zeros_u, zeros_v = np.zeros((37, 72)), np.zeros((36, 72))
    o = make_oracle(st, (zeros_u, zeros_v, np.full((37, 72), 5e4), np.zeros((37, 72))))
    o.prepare(1)
    o.space_operators(1, 1, 2, 1, 2, 900.0, 0)
    for name in ("du", "dv", "dgz"):
        assert np.all(o.internal(1, name, tend=True) == 0), name
    o.step(1800.0, 2)
    assert np.all(o.get(o.level, "u_lon") == 0) and np.all(o.get(o.level, "v_lat") == 0)
    assert np.all(o.get(o.level, "gz") == 5e4)


def test_rest_state_with_filters():
    """With the filters on, the filtered constant differs from the constant
    only by the rounding of the normalised weights, so the pressure gradient
    of a resting state is at rounding level."""
    st = _setup()
    o = make_oracle(st, (np.zeros((37, 72)), np.zeros((36, 72)), np.full((37, 72), 5e4), np.zeros((37, 72))))
    o.prepare(1)
    o.space_operators(1, 1, 2, 1, 2, 900.0, 0)
    pgf_scale = 5e4 / st.mesh["de_lat"][:36].min()
    for name in ("du", "dv", "dgz"):
        assert np.abs(o.internal(1, name, tend=True)).max() <= 1e-14 * pgf_scale, name


def _forward(st, ic, pass_=1, dt=900.0):
    o = make_oracle(st, ic)
    o.prepare(1)
    o.space_operators(1, 1, 2, 1, 2, dt, pass_)
    return o


def test_mass_flux_divergence_integrates_to_zero():
    st = _setup()
    o = _forward(st, seeded_ic(st, lmax=10, seed=5))
    dgz = o.internal(1, "dgz", tend=True)
    ac = st.mesh["area_cell"][:, None]
    terms = dgz * ac
    assert abs(terms.sum()) <= 1e-13 * np.abs(terms).sum()


def test_mass_conserved_over_steps():
    st = _setup(180, 91, 600.0)
    o = make_oracle(st, seeded_ic(st))
    o.prepare(1)
    tm0 = o.diagnose(1)[0]
    o.step(600.0, 30)
    tm = o.diagnose()[0]
    assert abs(tm - tm0) <= 1e-13 * tm0


def _ip(o, st, tend):
    nlat = st.nl.num_lat
    al = st.mesh["area_lon"][:, None]
    at = st.mesh["area_lat"][:nlat - 1, None]
    ac = st.mesh["area_cell"][:, None]
    mfx, mfy = o.internal(1, "mfx_lon")[1:-1], o.internal(1, "mfy_lat")
    g = lambda n: o.internal(1, n, tend=True)
    cf = [g("qhv")[1:-1] * mfx * al[1:-1], -g("qhu") * mfy * at]
    ke = [g("dkedlon")[1:-1] * mfx * al[1:-1] * 2, g("dkedlat") * mfy * at * 2,
          (g("dmfdlon") + g("dmfdlat")) * o.internal(1, "ke") * ac]
    pe = [g("pgf_lon")[1:-1] * mfx * al[1:-1] * 2, g("pgf_lat") * mfy * at * 2,
          (g("dmfdlon") + g("dmfdlat")) * o.internal(1, "gz") * ac]
    return [(sum(t.sum() for t in ts), sum(np.abs(t).sum() for t in ts)) for ts in (cf, ke, pe)]


@pytest.mark.parametrize("opts", [dict(), dict(pv_scheme="midpoint"), dict(upwind_order_pv=1)])
def test_energy_identities(opts):
    # filters off so that pgf is the gradient of gz itself (gz_f = gz)
    st = _setup(max_wave_speed=0.0, **opts)
    o = _forward(st, seeded_ic(st, lmax=10, seed=6), pass_=0)
    for name, (s, mag) in zip(("ip_cf", "ip_ke", "ip_pe"), _ip(o, st, None)):
        assert abs(s) <= 1e-13 * mag, (name, s, mag)


def test_energy_identity_detects_a_wrong_sign():
    """The identity check is sharp: flipping the sign of one pole term breaks it."""
    st = _setup(max_wave_speed=0.0)
    o = _forward(st, seeded_ic(st, lmax=10, seed=6), pass_=0)
    (s, mag) = _ip(o, st, None)[1]
    nlat = st.nl.num_lat
    ac = st.mesh["area_cell"]
    g = o.internal(1, "dmfdlat", tend=True)
    bad = s - 2 * (g[0] * o.internal(1, "ke")[0] * ac[0]).sum()
    assert abs(bad) > 1e-9 * mag


def test_hemispheric_mirror_symmetry():
    st = _setup(72, 37, 1800.0)
    u, v, gz, gzs = seeded_ic(st, lmax=8, seed=7)
    u = 0.5 * (u + u[::-1])
    v = 0.5 * (v - v[::-1])
    gz = 0.5 * (gz + gz[::-1])
    o = make_oracle(st, (u, v, gz, gzs))
    o.prepare(1)
    o.step(1800.0, 6)
    uu, vv, gg = (o.get(o.level, n) for n in ("u_lon", "v_lat", "gz"))
    assert np.abs(uu - uu[::-1]).max() <= 1e-11 * np.abs(uu).max()
    assert np.abs(vv + vv[::-1]).max() <= 1e-11 * np.abs(vv).max()
    assert np.abs(gg - gg[::-1]).max() <= 1e-12 * np.abs(gg).max()


def _williamson2_error(nlon, nlat, dt, steps):
    st = _setup(nlon, nlat, dt)
    ic = init_cond.steady_geostrophic(nlon, nlat, st.mesh.metrics, st.nl.omega, st.nl.radius)
    o = make_oracle(st, ic)
    o.prepare(1)
    o.step(dt, steps)
    gz = o.get(o.level, "gz")
    w = np.cos(st.mesh["full_lat"])[:, None] * np.ones((1, nlon))
    return np.sqrt(((gz - ic[2]) ** 2 * w).sum() / (ic[2] ** 2 * w).sum())


def test_steady_geostrophic_converges():
    e_coarse = _williamson2_error(72, 37, 1200.0, 72)   # 1 day
    e_fine = _williamson2_error(144, 73, 600.0, 144)    # 1 day
    assert e_coarse < 2e-3
    assert e_fine < 0.6 * e_coarse


def test_pc2_composition():
    st = _setup()
    ic = seeded_ic(st, lmax=8, seed=9)
    a, b = make_oracle(st, ic), make_oracle(st, ic)
    a.prepare(1)
    b.prepare(1)
    dt = st.nl.dt_dyn
    a.step(dt, 1)
    # time_schemes_mod.F90:244-258 with step_forward_backward (:102-118)
    for star, new, h in ((1, 2, dt / 2), (2, 3, dt / 2), (3, 2, dt)):
        b.space_operators(1, star, new, 1, 2, h, 1)
        b.update_state(1, 1, new, h, 1)
        b.space_operators(1, star, new, 2, 1, h, 2)
        b.update_state(2, 1, new, h, 2)
    for n in ("u_lon", "v_lat", "gz"):
        assert np.array_equal(a.get(a.level, n), b.get(2, n))


def test_zonal_sum_order():
    """gmo_zonal_sum adds 64 strided partials pairwise; pinned on values whose
    float sum depends on the order."""
    x = np.array([1e16, 1.0, -1e16, 1.0] * 40)
    p = [0.0] * 64
    for l in range(64):
        acc = 0.0
        for v in x[l::64]:
            acc += v
        p[l] = acc
    w = 1
    while w < 64:
        for l in range(0, 64, 2 * w):
            p[l] = p[l] + p[l + w]
        w *= 2
    assert oracle.zonal_sum(x) == p[0]
