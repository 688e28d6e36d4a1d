"""GPU parity of the shallow-water path against the oracle.

Bar: FP64 fields within 1e-12 relative per step (the kernels and the
oracle use the same operation order with no fused multiply-add, so most
comparisons are exact; the tolerance is the north-star bar).
"""
import numpy as np
import pytest

from helpers import STATE, make_gpu, make_oracle, namelist, rel_err, seeded_ic
from gmcore_amd.host import Setup

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _pair(nl, lmax=20, seed=0):
    st = Setup(nl)
    ic = seeded_ic(st, lmax=lmax, seed=seed)
    return st, make_oracle(st, ic), make_gpu(st, ic)


def test_prepare_matches():
    st, o, g = _pair(namelist())
    o.prepare(1)
    g.prepare(1)
    for f in ("u_f", "v_f", "gz_f", "pv", "ke", "m", "div"):
        assert rel_err(g.download(1, f), o.get(1, f)) <= TOL, f


@pytest.mark.parametrize("pass_,pv", [(0, "upwind"), (1, "upwind"), (1, "tvd")])
def test_forward_pass(pass_, pv):
    st, o, g = _pair(namelist(pv_scheme=pv))
    o.prepare(1)
    g.prepare(1)
    dt = 300.0
    o.space_operators(1, 1, 2, 1, 2, dt, pass_)
    g.space_operators(1, 1, 2, 1, 2, dt, pass_)
    for f in ("du", "dv", "dgz"):
        e = rel_err(g.download(1, f), o.get(1, f))
        assert e <= TOL, (f, e)


def test_update_state_and_backward():
    st, o, g = _pair(namelist())
    o.prepare(1)
    g.prepare(1)
    dt = 300.0
    for m in (o, g):
        m.space_operators(1, 1, 2, 1, 2, dt, 1)
        m.update_state(1, 1, 2, dt, 1)
        m.space_operators(1, 1, 2, 2, 1, dt, 2)
        m.update_state(2, 1, 2, dt, 2)
    for f in ("gz", "gz_f", "u_lon", "v_lat", "u_f", "v_f", "du", "dv"):
        lev = 2 if f in ("du", "dv") else 2
        e = rel_err(g.download(lev, f), o.get(lev, f))
        assert e <= TOL, (f, e)


@pytest.mark.parametrize("opts", [
    dict(),
    dict(pv_scheme="midpoint"),
    dict(upwind_order_pv=1),
    dict(coriolis_scheme=2),
    dict(use_div_damp=True),
    dict(use_div_damp=True, div_damp_order=4),
    dict(pv_scheme="tvd"),
    dict(pv_scheme="tvd", tvd_limiter_type="mc"),
    dict(pv_scheme="tvd", tvd_limiter_type="none", time_scheme="wrfrk3"),
    dict(time_scheme="wrfrk3"),
    dict(time_scheme="euler", dt_dyn=120.0),
    dict(max_wave_speed=0.0),
    dict(pv_pole_stokes=False),
])
def test_steps_match(opts):
    nl = namelist(**{k: v for k, v in opts.items() if k != "dt_dyn"})
    if "dt_dyn" in opts:
        nl.dt_dyn = opts["dt_dyn"]
    st, o, g = _pair(nl)
    o.prepare(1)
    g.prepare(1)
    for n in range(3):
        o.step(nl.dt_dyn, 1)
        g.step(1)
        assert g.level == o.level
        for f in STATE:
            e = rel_err(g.download(g.level, f), o.get(o.level, f))
            assert e <= TOL, (opts, n, f, e)


@pytest.mark.parametrize("scheme,dt", [("pc2", 600.0), ("wrfrk3", 600.0), ("euler", 120.0)])
def test_diagnose_reads_the_reference_fields(scheme, dt):
    """With gmc_keep_derived on, the stages leave each star state's
    forward-pass fields in its level as the reference's operators do, and
    gmc_diagnose gives the reference's tm, te, tpe and tav: te and tpe mix
    the final winds and masses with the mass fluxes and pv of the second
    stage's star state (gmcore_mod.F90:230-313).  The oracle's diagnose reads
    its stored fields the same way; the fresh sums differ from them by
    1e-7 .. 1e-5, far above the 1e-12 bar."""
    nl = namelist(time_scheme=scheme, dt=dt)
    st, o, g = _pair(nl)
    g.keep_derived(True)
    o.prepare(1)
    g.prepare(1)
    for n in range(3):
        if n:
            o.step(dt, 1)
            g.step(1)
        do, dg = o.diagnose(), g.diagnose()
        if n != 1 or scheme != "euler":   # after one euler step the reference reads a never-written level
            for k in range(4):
                assert abs(dg[k] - do[k]) <= 1e-12 * abs(do[k]), (n, k, dg[k], do[k])
    fresh = o.diagnose(fresh=True)
    assert abs(fresh[1] - do[1]) > 1e-9 * abs(do[1])


def test_diagnose_fresh_fields():
    """Without gmc_keep_derived the forward-pass fields are formed from the
    level's own state, as the oracle's fresh sums."""
    st, o, g = _pair(namelist())
    o.prepare(1)
    g.prepare(1)
    o.step(600.0, 2)
    g.step(2)
    do, dg = o.diagnose(fresh=True), g.diagnose()
    assert abs(dg[0] - do[0]) <= 1e-13 * abs(do[0])       # tm
    for k in (1, 2, 3):
        assert abs(dg[k] - do[k]) <= 1e-12 * abs(do[k]), (k, dg[k], do[k])


def test_long_step_runs_equal_single_steps():
    """gmc_step(ctx, dt, n) replays graphs of 8 steps while 8 or more remain,
    then graphs of one step: 19 steps in one call (8 + 8 + 3), and 8 after a
    single step (the block starting from the other level), give the same bits
    as the same steps taken one call at a time, with keep_derived on for the
    last steps so that both graph keys are used."""
    st = Setup(namelist())
    ic = seeded_ic(st)
    a, b = make_gpu(st, ic), make_gpu(st, ic)
    for m in (a, b):
        m.prepare(1)
    a.step(19)
    for _ in range(19):
        b.step(1)
    assert a.level == b.level
    for f in STATE:
        assert np.array_equal(a.download(a.level, f), b.download(b.level, f)), f
    for m in (a, b):
        m.step(1)
        m.keep_derived(True)
    a.step(8)
    for _ in range(8):
        b.step(1)
    assert a.level == b.level
    for f in STATE + ("u_f", "v_f", "gz_f"):
        assert np.array_equal(a.download(a.level, f), b.download(b.level, f)), f
    assert np.array_equal(a.download(3, "pv"), b.download(3, "pv"))   # the star level's kept pv
    assert np.array_equal(a.diagnose(), b.diagnose())
