"""The oracle's dynamics with tracers (time_integrate + adv_run(block, new),
gmcore_mod.F90:577-591).  Every adv_run keeps the dynamics' m as old_m
(copy_old_m, adv_mod.F90:309); an advancing call carries the tracer mass
from old_m to the new m by the flux form (adv_mod.F90:183-234), so
Σ m_new q_new A equals Σ old_m q_old A to rounding while the fixer does not
act, whatever mass fluxes the batch accumulated.  (The polar longitude
damping smooths q, not q m, so the check runs without it.)  The batch
advances at the time steps k * nstep, k >= 1."""
import numpy as np
import pytest

from coupled_helpers import coupled_oracle, initial, setups, tracer_mass


@pytest.mark.parametrize("nstep", [1, 2])
def test_each_advance_carries_the_tracer_mass(nstep):
    sw, adv = setups(nstep, ntracers=2, damp_passes=0)
    ic, qs = initial(sw, ntracers=3)
    qs = qs[:2]                               # the cosine-bell sums: the fixer never acts on them
    c = coupled_oracle(sw, adv, ic, qs)
    m_prev = c.sw.internal(c.sw.level, "m")
    advanced = []
    for n in range(6):
        before = [c.adv.tracer(c.adv.old, l) for l in range(len(qs))]
        old = c.adv.old
        assert c.step() == 0
        m = c.sw.internal(c.sw.level, "m")
        if c.adv.old != old:
            advanced.append(n)
            for l in range(len(qs)):
                want = tracer_mass(before[l], m_prev, sw)
                got = tracer_mass(c.adv.tracer(c.adv.old, l), m, sw)
                assert abs(got - want) <= 1e-13 * want, (n, l)
        else:
            for l in range(len(qs)):
                assert np.array_equal(c.adv.tracer(c.adv.old, l), before[l])
        m_prev = m
    assert c.adv.fixes == 0
    assert advanced == [t for t in range(6) if t > 0 and t % nstep == 0]
    for l in range(len(qs)):
        assert c.adv.tracer(c.adv.old, l).min() >= 0
