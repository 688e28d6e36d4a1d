"""Tracer advection inside the dynamics step on the GPU against the oracle:
after every gmc_step the batch runs adv_run(block, new) on the model's
state(new) through gmc_adv_run_dyn (u_f, v_f, the second stage's mfx_lon and
mfy_lat, m), with gmc_adv_init_dyn standing for gmcore_run's accumulation
before the loop (gmcore_mod.F90:117-131, 577-591).  Bar: every tracer
bit-identical to the coupled oracle after every step, the same fixer count;
PPM and van Leer fluxes, batches advanced every 1 and 2 dynamics steps, the
default three polar damping passes and a uniform-random tracer (which the
fixer works on at the two-step cadence)."""
import numpy as np
import pytest

from coupled_helpers import NLAT, NLON, coupled_oracle, initial, setups
from gmcore_amd.adv import Advection
from helpers import make_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nstep,flux", [(1, "ppm"), (2, "ppm"), (1, "van_leer")])
def test_coupled_tracers_match_oracle(nstep, flux):
    sw, adv = setups(nstep, ntracers=3, flux=flux)
    ic, qs = initial(sw, ntracers=3)
    c = coupled_oracle(sw, adv, ic, qs)
    g = make_gpu(sw, ic)
    g.keep_derived(True)
    g.prepare(1)
    a = Advection(adv)
    for l, q in enumerate(qs):
        a.set_tracer(1, l, q)
    a.init_dyn(g)
    for n in range(8):
        assert c.step() == 0
        g.step(1)
        a.run_dyn(g, n)
        assert a.old == c.adv.old, n
        for l in range(3):
            got, want = a.tracer(a.old, l), c.adv.tracer(c.adv.old, l)
            assert np.array_equal(got, want), (n, l, float(np.abs(got - want).max()))
    assert a.fixes == c.adv.fixes
    if nstep == 2:
        assert c.adv.fixes > 0                # the random tracer undershoots at this cadence
    a.close()
    g.close()


def test_run_dyn_needs_the_kept_mass_fluxes():
    sw, adv = setups(1)
    ic, qs = initial(sw)
    g = make_gpu(sw, ic)
    g.prepare(1)
    a = Advection(adv)
    a.init_dyn(g)
    g.step(1)                                 # without gmc_keep_derived
    with pytest.raises(Exception):
        a.run_dyn(g, 0)
    a.close()
    g.close()
