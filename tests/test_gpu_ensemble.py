"""Batched ensembles (include/gmcore_ens.h) against per-member plain runs.

Bar: every member of the ensemble holds, after every step, exactly the bits
gmc_step gives the same member run on its own, for every launch shape the
tuner can pick; the tuner leaves the state where it found it.
"""
import numpy as np
import pytest

from helpers import STATE
from gmcore_amd import init_cond
from gmcore_amd.host import Ensemble, Setup, ShallowWater
from gmcore_amd.namelist import Namelist

pytestmark = pytest.mark.gpu

# (nlon, nlat, dt) of the members: uneven sizes, each with its own dt
MEMBERS = [(72, 37, 1800.0)] * 3 + [(144, 73, 900.0)] * 2 + [(180, 91, 600.0)]


def _member(nlon, nlat, dt, k, scheme="pc2"):
    st = Setup(Namelist(num_lon=nlon, num_lat=nlat, dt_dyn=dt, time_scheme=scheme))
    ic = init_cond.seeded_state(nlon, nlat, st.mesh.metrics, 8, 0, perturb=(0.01, 100 + k),
                                omega=st.nl.omega, radius=st.nl.radius)
    m = ShallowWater(st)
    m.set_state(*ic)
    m.prepare(1)
    return m


def _build(scheme="pc2"):
    return [_member(*spec, k, scheme) for k, spec in enumerate(MEMBERS)]


def _assert_same(ens_members, plain_members, what):
    for k, (a, b) in enumerate(zip(ens_members, plain_members)):
        assert a.level == b.level, (what, k)
        sa, sb = a.state(), b.state()
        for f in STATE:
            assert np.array_equal(sa[f], sb[f]), (what, k, f)


@pytest.mark.parametrize("scheme", ["pc2", "wrfrk3"])
def test_ensemble_equals_plain_runs(scheme):
    ens_m, plain = _build(scheme), _build(scheme)
    ens = Ensemble(ens_m)
    for n in range(3):
        ens.step(1)
        for m in plain:
            m.step(1)
        _assert_same(ens_m, plain, f"step {n}")
    ens.step(2)
    for m in plain:
        m.step(2)
    _assert_same(ens_m, plain, "step 5")
    # long runs replay graphs of 8 steps: from either parity, and the
    # members' own gmc_step runs in graphs of 8 steps as well
    for n in (10, 9):
        ens.step(n)
        for m in plain:
            for _ in range(n):
                m.step(1)
        _assert_same(ens_m, plain, f"{n} more steps")
    ens.close()


def test_every_shape_gives_the_same_bits():
    ref = _build()
    for m in ref:
        m.step(2)
    ncand = next(k for k in range(64) if not Ensemble.shape_name(k))
    assert ncand >= 6
    for shape in range(ncand):
        ms = _build()
        ens = Ensemble(ms)
        ens.set_shape(shape)
        ens.step(2)
        _assert_same(ms, ref, f"shape {shape} ({Ensemble.shape_name(shape)})")
        ens.close()


def test_autotune_keeps_state():
    ens_m, plain = _build(), _build()
    ens = Ensemble(ens_m)
    choice, times = ens.autotune(trial_steps=2)
    assert 0 <= choice < len(times) and np.all(times > 0)
    ens.step(3)
    for m in plain:
        m.step(3)
    _assert_same(ens_m, plain, f"after autotune (shape {choice})")
    ens.close()
