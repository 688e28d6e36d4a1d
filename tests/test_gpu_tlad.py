"""Tangent-linear and adjoint of the shallow-water step (include/gmcore_tlad.h).

- dot-product test: <M dx, l> = <dx, M^T l> over several pc2 steps, to 1e-12
  relative, with filters, upwind-3 pv, Stokes pole vorticity and the pole
  smoothing in the step;
- the TL against centred finite differences of the nonlinear step;
- the gradient of a cost (observed gz at the window's end plus a background
  term) against centred finite differences of that cost;
- the trajectory: gmc_tl_step advances the state exactly as gmc_step, and
  gmc_ad_step / gmc_gradient leave it where it was;
- checkpointing: every checkpoint interval gives the same bits;
- the dot-product test at 360 x 181 (the adjoint's row kernels cut polar
  rows into column chunks) and with Coriolis scheme 2.
"""
import numpy as np
import pytest

from gmcore_amd import init_cond
from gmcore_amd.host import TLAD, Setup, ShallowWater
from gmcore_amd.namelist import Namelist

pytestmark = pytest.mark.gpu

NLON, NLAT, DT = 72, 37, 1800.0


def _model(scheme="pc2", pv="upwind", ic_seed=11):
    st = Setup(Namelist(num_lon=NLON, num_lat=NLAT, dt_dyn=DT, time_scheme=scheme, pv_scheme=pv))
    ic = init_cond.seeded_state(NLON, NLAT, st.mesh.metrics, 8, ic_seed, omega=st.nl.omega, radius=st.nl.radius)
    m = ShallowWater(st)
    m.set_state(*ic)
    m.prepare(1)
    return m, ic


def _rand(seed, scale=(0.5, 0.5, 100.0)):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((NLAT, NLON)) * scale[0]
    u[[0, -1]] = 0.0
    v = rng.standard_normal((NLAT - 1, NLON)) * scale[1]
    gz = rng.standard_normal((NLAT, NLON)) * scale[2]
    return u, v, gz


def _dot(a, b):
    return sum(float(np.sum(x[1:-1] * y[1:-1])) if k == 0 else float(np.sum(x * y))
               for k, (x, y) in enumerate(zip(a, b)))


def _tl(m, dx, n):
    t = TLAD(m)
    lv = m.level
    for name, a in zip(("u", "v", "gz"), dx):
        t.set(lv, name, a)
    t.tl_step(n)
    out = tuple(t.get(m.level, name) for name in ("u", "v", "gz"))
    t.close()
    return out


def _ad(m, lam, n):
    t = TLAD(m)
    lv = m.level
    for name, a in zip(("u", "v", "gz"), lam):
        t.set(lv, name, a)
    t.ad_step(n)
    out = tuple(t.get(lv, name) for name in ("u", "v", "gz"))
    t.close()
    return out


@pytest.mark.parametrize("scheme,pv", [("pc2", "upwind"), ("pc2", "midpoint"), ("wrfrk3", "upwind")])
def test_dot_product(scheme, pv):
    n = 3
    dx = _rand(1)
    lam = _rand(2, scale=(1.0, 1.0, 1e-2))
    m1, _ = _model(scheme, pv)
    mdx = _tl(m1, dx, n)
    m2, _ = _model(scheme, pv)
    mtl = _ad(m2, lam, n)
    lhs, rhs = _dot(mdx, lam), _dot(dx, mtl)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def test_tl_trajectory_equals_gmc_step():
    m1, _ = _model()
    _tl(m1, _rand(3), 4)
    m2, _ = _model()
    m2.step(4)
    assert m1.level == m2.level
    s1, s2 = m1.state(), m2.state()
    for f in s1:
        assert np.array_equal(s1[f], s2[f]), f


def test_adjoint_leaves_state():
    m, ic = _model()
    before = m.state()
    _ad(m, _rand(4), 3)
    after = m.state()
    for f in before:
        assert np.array_equal(before[f], after[f]), f


def test_tl_against_finite_differences():
    n = 3
    dx = _rand(5)
    m0, ic = _model()
    mdx = _tl(m0, dx, n)
    eps = 1e-3
    outs = []
    for sgn in (1.0, -1.0):
        m, _ = _model()
        m.set_state(ic[0] + sgn * eps * dx[0], ic[1] + sgn * eps * dx[1], ic[2] + sgn * eps * dx[2], ic[3])
        m.prepare(1)
        m.step(n)
        s = m.state()
        outs.append((s["u_lon"], s["v_lat"], s["gz"]))
    for k, name in enumerate(("u", "v", "gz")):
        fd = (outs[0][k] - outs[1][k]) / (2 * eps)
        rows = slice(1, -1) if k == 0 else slice(None)
        err = np.abs(fd[rows] - mdx[k][rows]).max() / np.abs(mdx[k][rows]).max()
        assert err <= 1e-5, (name, err)


def _obs(seed, nobs=200):
    rng = np.random.default_rng(seed)
    cells = rng.choice(NLAT * NLON, size=nobs, replace=False)
    return cells


def _cost_and_grad(m, cells, y, xb, n, ckpt):
    t = TLAD(m)
    J, g = t.gradient(n, ckpt, cells, y, 50.0, xb, 10.0)
    t.close()
    return J, g


def test_gradient_against_finite_differences():
    n, ckpt = 6, 2
    m, ic = _model()
    cells = _obs(7)
    truth, _ = _model(ic_seed=12)
    truth.step(n)
    y = truth.state()["gz"].ravel()[cells] + np.random.default_rng(8).normal(0, 50.0, len(cells))
    xb = (ic[0] + _rand(9, (0.2, 0.2, 20.0))[0], ic[1], ic[2])
    J, g = _cost_and_grad(m, cells, y, xb, n, ckpt)
    d = _rand(10)
    dJ = _dot(g, d)
    eps = 1e-3
    Js = []
    for sgn in (1.0, -1.0):
        mm, _ = _model()
        mm.set_state(ic[0] + sgn * eps * d[0], ic[1] + sgn * eps * d[1], ic[2] + sgn * eps * d[2], ic[3])
        mm.prepare(1)
        Js.append(_cost_and_grad(mm, cells, y, xb, n, ckpt)[0])
    fd = (Js[0] - Js[1]) / (2 * eps)
    assert abs(fd - dJ) <= 1e-6 * abs(dJ), (fd, dJ)
    # the state is left at x0
    assert np.array_equal(m.state()["gz"], ic[2])


def test_checkpoint_interval_gives_same_bits():
    n = 6
    cells = _obs(11)
    res = []
    for ckpt in (1, 4, 6):
        m, ic = _model()
        y = np.full(len(cells), 5.0e4)
        res.append(_cost_and_grad(m, cells, y, (ic[0], ic[1], ic[2]), n, ckpt))
    for J, g in res[1:]:
        assert J == res[0][0]
        for a, b in zip(g, res[0][1]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("scheme,pv", [("pc2", "upwind"), ("wrfrk3", "midpoint")])
def test_gpu_adjoint_matches_cpu_adjoint(scheme, pv):
    """The FP64 CPU adjoint (oracle/adjoint.py, numpy on the oracle's
    trajectory) and the GPU adjoint agree to 1e-12 relative."""
    import oracle
    from oracle.adjoint import StepAdjoint
    from helpers import make_oracle
    n = 2
    lam = _rand(2, scale=(1.0, 1.0, 1e-2))
    m, ic = _model(scheme, pv)
    g = _ad(m, lam, n)
    o = make_oracle(m.setup, ic)
    o.prepare(1)
    x0 = {k: o.get(1, k) for k in ("u_lon", "v_lat", "gz", "u_f", "v_f", "gz_f")}
    c = StepAdjoint(o, m.setup).ad_steps(x0, *lam, DT, n)
    for k, name in enumerate(("u", "v", "gz")):
        a, b = (g[k][1:-1], c[k][1:-1]) if k == 0 else (g[k], c[k])
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), name


def _model_at(nlon, nlat, dt, coriolis):
    st = Setup(Namelist(num_lon=nlon, num_lat=nlat, dt_dyn=dt, time_scheme="pc2", coriolis_scheme=coriolis))
    ic = init_cond.seeded_state(nlon, nlat, st.mesh.metrics, 8, 11, omega=st.nl.omega, radius=st.nl.radius)
    m = ShallowWater(st)
    m.set_state(*ic)
    m.prepare(1)
    return m


def _rand_at(nlon, nlat, seed, scale):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((nlat, nlon)) * scale[0]
    u[[0, -1]] = 0.0
    return u, rng.standard_normal((nlat - 1, nlon)) * scale[1], rng.standard_normal((nlat, nlon)) * scale[2]


@pytest.mark.parametrize("nlon,nlat,dt,coriolis", [(360, 181, 300.0, 1), (72, 37, 1800.0, 2)])
def test_dot_product_chunked_rows_and_coriolis2(nlon, nlat, dt, coriolis):
    """At 360 x 181 the adjoint's row kernels cut the polar rows into
    column chunks that read each other's columns as halo; with Coriolis
    scheme 2 the taping stages also keep the tangential mass fluxes, which
    the adjoint of that scheme reads."""
    n = 2
    dx = _rand_at(nlon, nlat, 1, (0.5, 0.5, 100.0))
    lam = _rand_at(nlon, nlat, 2, (1.0, 1.0, 1e-2))
    mdx = _tl(_model_at(nlon, nlat, dt, coriolis), dx, n)
    mtl = _ad(_model_at(nlon, nlat, dt, coriolis), lam, n)
    lhs, rhs = _dot(mdx, lam), _dot(dx, mtl)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs)), (lhs, rhs)
