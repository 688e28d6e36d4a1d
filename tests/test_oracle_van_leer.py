"""Pins of the oracle's van Leer flux (hflx_van_leer_cell, ffsl_mod.F90:
556-616), selected by ffsl_flux_type = 'van_leer'.

- 1-D reduction: in solid-body zonal flow (v = 0) a mid-latitude row
  advances by the x-flux alone, and with the unlimited slope ('none') the
  flux of cell i is u (q_i + (1 - c)/4 (q_{i+1} - q_{i-1})) for 0 < c < 1:
  Fromm's scheme.  A zonal Fourier mode is then multiplied by its closed-form
  amplification factor 1 - g [(1 - e^{-it}) + (1 - c)/4 (e^{it} - 1 - e^{-it}
  + e^{-2it})], g = dt u le_lon / A_cell, c the row's Courant number.  A
  wrong sign of the slope term, a shifted upwind cell or a missing (1 - c)
  fails it.
- a uniform tracer stays uniform in a nondivergent deformational flow, mass
  is conserved to rounding and the monotone limiter keeps q >= 0;
- the tangent linear and adjoint refuse the van Leer flux.
"""
import cmath
import math

import numpy as np
import pytest

import oracle
from gmcore_amd import init_cond
from gmcore_amd.adv import AdvSetup
from gmcore_amd.namelist import Namelist

NLON, NLAT, DT = 72, 37, 3600.0


def _setup(limiter="none", ntracers=2, hw=20):
    nl = Namelist(num_lon=NLON, num_lat=NLAT, dt_adv=DT, dt_dyn=DT, limiter_type=limiter,
                  ffsl_flux_type="van_leer")
    return AdvSetup(nl, ntracers, lon_halo_width=hw)


def _drive(st, qs, u, v, nsteps):
    o = oracle.AdvOracle(st.mesh_struct, st.cfg)
    for l, q in enumerate(qs):
        o.set_tracer(1, l, q)
    ones = np.ones((NLAT, NLON))
    for n in range(nsteps):
        o.set_wind(u, v, u, v, ones)
        assert o.run(n) == 0
    return o


@pytest.mark.parametrize("k", [1, 5, 12])
def test_zonal_mode_amplification(k):
    st = _setup()
    M = st.mesh
    U = 30.0
    u = U * np.cos(M["full_lat"])[:, None] * np.ones((1, NLON))
    u[[0, -1]] = 0.0
    v = np.zeros((NLAT - 1, NLON))
    lam = (np.arange(NLON) + 0.5) * M.dlon - 0.5 * M.dlon    # cell centres
    q = 2.0 + np.cos(k * lam)[None, :] * np.ones((NLAT, 1))
    o = _drive(st, [q, np.ones((NLAT, NLON))], u, v, 2)      # step 0 only accumulates
    r = 9                                                     # a mid-latitude row, no lon damping
    assert st.wgt[r] == 0.0
    c = o.internal("cflx")[r]
    assert np.ptp(c) < 1e-12 and 0 < c[0] < 1
    c = float(c[0])
    g = DT * u[r, 0] * M["le_lon"][r] / M["area_cell"][r]
    t = k * M.dlon
    G = 1 - g * ((1 - cmath.exp(-1j * t)) + (1 - c) / 4 * (cmath.exp(1j * t) - 1 - cmath.exp(-1j * t)
                                                            + cmath.exp(-2j * t)))
    want = 2.0 + np.real(G * np.exp(1j * k * lam))
    got = o.tracer(o.old, 0)[r]
    assert abs(G - 1) > 1e-3          # the step moves the mode (phase and amplitude)
    assert np.abs(got - want).max() <= 1e-12, np.abs(got - want).max()
    # the uniform tracer stays uniform
    assert np.abs(o.tracer(o.old, 1) - 1.0).max() <= 1e-13


def test_deformational_flow_invariants():
    st = _setup(limiter="mono", ntracers=2)
    ub, vb = init_cond.deform_wind_basis(NLON, NLAT, st.mesh.metrics, 6, 2, vmax=50.0)
    q0 = init_cond.cosine_bells(NLON, NLAT, st.mesh.metrics, 21)   # smooth: the fixer never acts
    ones = np.ones((NLAT, NLON))
    o = _drive(st, [q0, ones], ub * 0.3, vb * 0.3, 6)
    ac = st.mesh["area_cell"][:, None]
    m0 = math.fsum((q0 * ac).ravel())
    q = o.tracer(o.old, 0)
    assert o.fixes == 0
    assert abs(math.fsum((q * ac).ravel()) - m0) <= 1e-13 * m0
    assert q.min() >= 0.0
    assert np.abs(o.tracer(o.old, 1) - 1.0).max() <= 1e-12


def test_tangent_linear_refuses_van_leer():
    st = _setup()
    o = oracle.AdvOracle(st.mesh_struct, st.cfg)
    dq = np.zeros((2, NLAT, NLON))
    assert o.tlad_run(1, dq, None) != 0
