"""The adjoint and ensemble workloads at their full sizes on the GPU.

- Gradient window: 360 x 181, pc2, dt = 300 s, 48 steps, checkpoint every
  8 steps with recompute, 5000 observed gz cells.  The dot-product test of
  the TL and AD over the whole window (1e-12 relative), the gradient of the
  cost against a centred difference of the cost, and the same bits for two
  checkpoint intervals.
- Ensemble: 64 members of uneven size (32 x 180x91 at dt 600 s, 24 x 360x181
  at 300 s, 8 x 720x361 at 150 s), each with its own 1 %-perturbed seeded
  state, stepped in one batch; every member bit-identical to its plain
  gmc_step run, and one member of each size within 1e-12 of the oracle.
"""
import sys

import numpy as np
import pytest

from helpers import STATE, make_oracle, rel_err
from gmcore_amd import init_cond
from gmcore_amd.host import TLAD, Ensemble, Setup, ShallowWater
from gmcore_amd.namelist import Namelist

pytestmark = pytest.mark.gpu

NLON, NLAT, DT, WINDOW, CKPT, NOBS = 360, 181, 300.0, 48, 8, 5000
SIGMA_O, SIGMA_B = 50.0, 1.0


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def window():
    """Setup, background state, observation cells and values of the window."""
    st = Setup(Namelist(num_lon=NLON, num_lat=NLAT, dt_dyn=DT, time_scheme="pc2"))
    m_ = st.mesh.metrics
    truth_ic = init_cond.seeded_state(NLON, NLAT, m_, 30, 4, omega=st.nl.omega, radius=st.nl.radius)
    bg = init_cond.seeded_state(NLON, NLAT, m_, 30, 4, perturb=(0.01, 6), omega=st.nl.omega,
                                radius=st.nl.radius)
    truth = ShallowWater(st)
    truth.set_state(*truth_ic)
    truth.prepare(1)
    truth.step(WINDOW)
    rng = np.random.default_rng(5)
    cells = rng.choice(NLAT * NLON, size=NOBS, replace=False)
    y = truth.state()["gz"].ravel()[cells] + rng.normal(0.0, SIGMA_O, NOBS)
    truth.close()
    return st, bg, cells, y


def _model(st, x):
    m = ShallowWater(st)
    m.set_state(*x)
    m.prepare(1)
    return m


def _rand(seed, scale):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((NLAT, NLON)) * scale[0]
    u[[0, -1]] = 0.0
    return u, rng.standard_normal((NLAT - 1, NLON)) * scale[1], rng.standard_normal((NLAT, NLON)) * scale[2]


def _dot(a, b):
    return float(np.sum(a[0][1:-1] * b[0][1:-1]) + np.sum(a[1] * b[1]) + np.sum(a[2] * b[2]))


def _gradient(st, x, cells, y, xb, ckpt):
    m = _model(st, x)
    t = TLAD(m)
    J, g = t.gradient(WINDOW, ckpt, cells, y, SIGMA_O, xb, SIGMA_B)
    t.close()
    m.close()
    return J, g


def test_dot_product_over_the_window(window):
    st, bg, _, _ = window
    _log(f"[adj] dot product over {WINDOW} steps at {NLON}x{NLAT}")
    names = ("u", "v", "gz")
    dx = _rand(1, (0.5, 0.5, 100.0))
    lam = _rand(2, (1.0, 1.0, 1e-2))
    m = _model(st, bg)
    t = TLAD(m)
    lv = m.level
    for k, a in zip(names, dx):
        t.set(lv, k, a)
    t.tl_step(WINDOW)
    mdx = [t.get(m.level, k) for k in names]
    t.close()
    m.close()
    m = _model(st, bg)
    t = TLAD(m)
    lv = m.level
    for k, a in zip(names, lam):
        t.set(lv, k, a)
    t.ad_step(WINDOW)
    mtl = [t.get(lv, k) for k in names]
    t.close()
    m.close()
    lhs, rhs = _dot(mdx, lam), _dot(dx, mtl)
    _log(f"[adj] <M dx, l> = {lhs!r}, <dx, M^T l> = {rhs!r}")
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs)), (lhs, rhs)


@pytest.mark.parametrize("pv", ["upwind", "midpoint"])
def test_gradient_against_centred_difference(window, pv):
    """The upwind pv edge value has kinks (sign(v), abs(v) at v = 0,
    upwind_mod.F90:27-39, operators_mod.F90:738-795): a centred difference
    that straddles one converges only linearly, so over 48 steps at 65,160
    cells its error at eps = 1e-3 is a few 1e-6 and falls by about half per
    halving.  With the midpoint pv the step is smooth: at eps = 1e-3 the
    difference already sits on the round-off floor of the perturbed
    trajectories (gz rounded to its last bit in 5000 misfits, about 2e-9 of
    dJ), and both it and its Richardson extrapolate meet the gradient to
    1e-8."""
    st0, bg, cells, y = window
    st = st0 if pv == "upwind" else Setup(Namelist(num_lon=NLON, num_lat=NLAT, dt_dyn=DT, time_scheme="pc2",
                                                   pv_scheme=pv))
    xb = (bg[0], bg[1], bg[2])
    x0 = (bg[0] + _rand(9, (0.2, 0.2, 20.0))[0], bg[1], bg[2], bg[3])
    _log(f"[adj/{pv}] gradient over {WINDOW} steps, checkpoint every {CKPT}, {NOBS} observations")
    J, g = _gradient(st, x0, cells, y, xb, CKPT)
    d = _rand(10, (0.5, 0.5, 100.0))
    dJ = _dot(g, d)

    def centred(eps):
        Js = []
        for sgn in (1.0, -1.0):
            xs = tuple(x0[k] + sgn * eps * d[k] for k in range(3)) + (x0[3],)
            Js.append(_gradient(st, xs, cells, y, xb, CKPT)[0])
        return (Js[0] - Js[1]) / (2 * eps)

    fd1, fd2 = centred(1e-3), centred(5e-4)
    fdr = (4.0 * fd2 - fd1) / 3.0
    _log(f"[adj/{pv}] J = {J!r}, <grad, d> = {dJ!r}, centred differences {fd1!r} (1e-3), {fd2!r} (5e-4), "
         f"extrapolated {fdr!r}")
    assert abs(fd1 - dJ) <= 1e-5 * abs(dJ), (fd1, dJ)
    if pv == "upwind":
        assert abs(fd2 - dJ) <= 0.8 * abs(fd1 - dJ), (fd1, fd2, dJ)
    else:
        assert abs(fd1 - dJ) <= 1e-8 * abs(dJ), (fd1, dJ)
        assert abs(fdr - dJ) <= 1e-8 * abs(dJ), (fdr, dJ)


def test_checkpoint_intervals_give_the_same_gradient(window):
    st, bg, cells, y = window
    xb = (bg[0], bg[1], bg[2])
    J8, g8 = _gradient(st, bg, cells, y, xb, 8)
    J16, g16 = _gradient(st, bg, cells, y, xb, 16)
    assert J8 == J16
    for a, b in zip(g8, g16):
        assert np.array_equal(a, b)


# ---------------------------------------------------------------------------
ENS64 = [(180, 91, 600.0, 32), (360, 181, 300.0, 24), (720, 361, 150.0, 8)]


def test_ens64_equals_plain_runs_and_oracle():
    setups, ics, members = [], [], []
    k = 0
    for nlon, nlat, dt, count in ENS64:
        st = Setup(Namelist(num_lon=nlon, num_lat=nlat, dt_dyn=dt, time_scheme="pc2"))
        for _ in range(count):
            ic = init_cond.seeded_state(nlon, nlat, st.mesh.metrics, 20, 0, perturb=(0.01, 100 + k),
                                        omega=st.nl.omega, radius=st.nl.radius)
            m = ShallowWater(st)
            m.set_state(*ic)
            m.prepare(1)
            setups.append(st)
            ics.append(ic)
            members.append(m)
            k += 1
    _log(f"[ens64] {len(members)} members set up; 2 batched steps")
    ens = Ensemble(members)
    ens.step(2)
    ens.close()
    first = {}
    for k, (st, ic, a) in enumerate(zip(setups, ics, members)):
        b = ShallowWater(st)
        b.set_state(*ic)
        b.prepare(1)
        b.step(2)
        assert a.level == b.level, k
        sa, sb = a.state(), b.state()
        for f in STATE:
            assert np.array_equal(sa[f], sb[f]), (k, f)
        b.close()
        first.setdefault(st.nl.num_lon, k)
    _log(f"[ens64] every member equals its plain run; oracle check of members {sorted(first.values())}")
    for k in first.values():
        st, ic, a = setups[k], ics[k], members[k]
        o = make_oracle(st, ic)
        o.prepare(1)
        o.step(st.nl.dt_dyn, 2)
        for f in STATE:
            e = rel_err(a.download(a.level, f), o.get(o.level, f))
            assert e <= 1e-12, (k, f, e)
        o.close()
    for m in members:
        m.close()
