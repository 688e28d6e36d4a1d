"""The CPU adjoint (oracle/adjoint.py) against finite differences of the
oracle's own step, on the CPU: for J(x) = <l, M(x)> after n steps,
<dJ/dx, d> must equal (J(x + e d) - J(x - e d)) / (2 e)."""
import numpy as np
import pytest

import oracle
from oracle.adjoint import StepAdjoint

from helpers import make_oracle, namelist, seeded_ic
from gmcore_amd.host import Setup


def _rand(rng, nlat, nlon, scale):
    u = rng.standard_normal((nlat, nlon)) * scale[0]
    u[[0, -1]] = 0.0
    return u, rng.standard_normal((nlat - 1, nlon)) * scale[1], rng.standard_normal((nlat, nlon)) * scale[2]


def _final(st, ic, n):
    o = make_oracle(st, ic)
    o.prepare(1)
    o.step(st.nl.dt_dyn, n)
    return o.get(o.level, "u_lon"), o.get(o.level, "v_lat"), o.get(o.level, "gz")


def _dot(a, b):
    return float(np.sum(a[0][1:-1] * b[0][1:-1])) + float(np.sum(a[1] * b[1])) + float(np.sum(a[2] * b[2]))


@pytest.mark.parametrize("scheme,pv", [("pc2", "upwind"), ("wrfrk3", "midpoint")])
def test_cpu_adjoint_against_finite_differences(scheme, pv):
    nlon, nlat, n = 36, 19, 2
    st = Setup(namelist(nlon, nlat, 3600.0, time_scheme=scheme, pv_scheme=pv))
    ic = seeded_ic(st, lmax=6, seed=3)
    rng = np.random.default_rng(4)
    lam = _rand(rng, nlat, nlon, (1.0, 1.0, 1e-2))
    d = _rand(rng, nlat, nlon, (0.5, 0.5, 100.0))
    o = make_oracle(st, ic)
    o.prepare(1)
    x0 = {k: o.get(1, k) for k in ("u_lon", "v_lat", "gz", "u_f", "v_f", "gz_f")}
    grad = StepAdjoint(o, st).ad_steps(x0, *lam, st.nl.dt_dyn, n)
    eps = 1e-3
    J = []
    for sgn in (1.0, -1.0):
        icp = (ic[0] + sgn * eps * d[0], ic[1] + sgn * eps * d[1], ic[2] + sgn * eps * d[2], ic[3])
        J.append(_dot(lam, _final(st, icp, n)))
    fd = (J[0] - J[1]) / (2 * eps)
    ad = _dot(grad, d)
    # 4e-10 here; a 1 % error in one upwind-weight derivative gives 5e-7
    assert abs(fd - ad) <= 1e-8 * abs(ad), (fd, ad)
