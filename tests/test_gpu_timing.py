"""The per-kernel timing entry points compute what the plain entry points do.

gmc_kernel_times and gmc_adv_kernel_times launch each kernel with a start
and an end event stamped by its dispatch (hipExtLaunchKernelGGL), which
passes the kernel arguments' bytes as given; the plain paths (gmc_step from
a captured graph, gmc_adv_drive) launch the same kernels normally.  Bar: the
states after the same steps are bit-identical, and every class's mean
duration is positive and finite.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import STATE, make_gpu, namelist, seeded_ic
from gmcore_amd import init_cond
from gmcore_amd.adv import Advection, AdvSetup
from gmcore_amd.capi import dptr
from gmcore_amd.host import Setup
from gmcore_amd.namelist import Namelist

pytestmark = pytest.mark.gpu


def test_kernel_times_steps_like_gmc_step():
    # 360 x 181: polar rows cut into several filter chunks
    nl = namelist(360, 181, 300.0)
    setup = Setup(nl)
    ic = seeded_ic(setup, lmax=30, seed=4)
    a, b = make_gpu(setup, ic), make_gpu(setup, ic)
    for m in (a, b):
        m.prepare(1)
    a.step(3)
    kt = np.zeros(8)
    nk = C.c_int32()
    assert b.lib.gmc_kernel_times(b.ctx, nl.dt_dyn, 3, dptr(kt), 8, C.byref(nk)) == 0
    assert nk.value == 3
    assert [b.lib.gmc_kernel_name(k).decode() for k in range(3)] == ["k_tend", "k_gzf", "k_uvf"]
    assert np.all(np.isfinite(kt[:3])) and np.all(kt[:3] > 0) and np.all(kt[:3] < 100.0)   # ms
    assert a.level == b.level
    sa, sb = a.state(), b.state()
    for k in STATE:
        assert np.array_equal(sa[k], sb[k]), k
    # and the context keeps stepping from there like the other one
    a.step(1)
    b.step(1)
    sa, sb = a.state(), b.state()
    for k in STATE:
        assert np.array_equal(sa[k], sb[k]), k


def test_adv_kernel_times_steps_like_drive():
    nlon, nlat, dt, ntr = 180, 91, 1800.0, 3
    nl = Namelist(num_lon=nlon, num_lat=nlat, dt_adv=dt, dt_dyn=dt)
    st = AdvSetup(nl, ntr, lon_halo_width=60)
    m = st.mesh.metrics
    ub, vb = init_cond.deform_wind_basis(nlon, nlat, m, 10, 8, 50.0)
    qs = init_cond.tracer_set(nlon, nlat, m, 4, ntr)
    a, b = Advection(st), Advection(st)
    for x in (a, b):
        for l, q in enumerate(qs):
            x.set_tracer(1, l, q)
        x.set_wind_basis(ub, vb)
    sc = init_cond.wind_scales(6, dt)
    a.drive(0, sc)
    kt = np.zeros(8)
    nk = C.c_int32()
    assert b.lib.gmc_adv_kernel_times(b.h, 0, dptr(np.ascontiguousarray(sc)), 6, dptr(kt), 8, C.byref(nk)) == 0
    names = [b.lib.gmc_adv_kernel_name(k).decode() for k in range(nk.value)]
    assert names == ["k_adv_accum", "k_adv_cfl", "k_adv_inner", "k_adv_flux", "k_adv_fix"]
    assert np.all(np.isfinite(kt[:nk.value])) and np.all(kt[:nk.value] > 0) and np.all(kt[:nk.value] < 100.0)
    assert a.old == b.old
    for l in range(ntr):
        assert np.array_equal(a.tracer(a.old, l), b.tracer(b.old, l)), l
