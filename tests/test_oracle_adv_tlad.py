"""Pins of the oracle's tangent linear and adjoint of FFSL advection
(oracle/adv.c, gmo_adv_tlad_run) on the CPU.

With the winds held, one adv_run step is piecewise linear and positively
homogeneous in the tracers (limiters, PPM edge clipping and the positivity
fixer only choose among linear expressions), so:
- Euler's identity: the tangent at q applied to q itself is the step's
  output F(q), to rounding, in every piece the base selects (a dropped
  term, a wrong sign or index in one branch breaks it);
- central differences of the forward oracle, (F(q + e d) - F(q - e d)) / 2e,
  equal the tangent when no branch changes within +-e (exact for a linear
  piece up to rounding);
- the adjoint is the transpose of the tangent: <A^T l, d> = <l, A d> over a
  window of steps;
- ppm's Jacobian: Euler's identity and central differences per stencil.
"""
import numpy as np
import pytest

import oracle
from gmcore_amd import init_cond
from gmcore_amd.adv import AdvSetup
from gmcore_amd.capi import LIMITERS
from gmcore_amd.namelist import Namelist

NLAT, NLON = 37, 72


def _setup(limiter="mono", dt=3600.0, ntracers=2):
    nl = Namelist(num_lon=NLON, num_lat=NLAT, dt_adv=dt, dt_dyn=dt, limiter_type=limiter)
    return AdvSetup(nl, ntracers, lon_halo_width=20, lon_damp_passes=3)


def _fields(st, kind, seed=1):
    rng = np.random.default_rng(seed)
    nt = st.cfg.ntracers
    if kind == "smooth":
        lat = np.linspace(-90, 90, NLAT)[:, None] * np.pi / 180
        lon = np.arange(NLON)[None, :] * 2 * np.pi / NLON
        q = np.stack([1.0 + 0.5 * np.cos(lat) ** 2 * np.sin(2 * lon + l) + 0.05 * rng.standard_normal((NLAT, NLON))
                      for l in range(nt)])
    elif kind == "square":
        q = np.zeros((nt, NLAT, NLON))
        q[:, 12:25, 20:40] = 1.0
        q += 1e-3 * rng.uniform(0, 1, q.shape)
    else:
        q = rng.uniform(0, 1, (nt, NLAT, NLON))
    return np.ascontiguousarray(q)


class Window:
    """A prescribed-flow window: wind basis times a scale per step, m = 1."""

    def __init__(self, st, q0, nsteps, seed=3):
        self.st = st
        self.ub, self.vb = init_cond.deform_wind_basis(NLON, NLAT, st.mesh.metrics, 6, seed, 40.0)
        self.sc = init_cond.wind_scales(nsteps, st.nl.dt_adv)
        self.nsteps = nsteps
        self.o = oracle.AdvOracle(st.mesh_struct, st.cfg)
        for l in range(st.cfg.ntracers):
            self.o.set_tracer(1, l, q0[l])
            self.o.set_tracer(2, l, q0[l])

    def wind(self, o, n):
        u, v = self.ub * self.sc[n], self.vb * self.sc[n]
        o.set_wind(u, v, u, v, np.ones((NLAT, NLON)))

    def out(self, o):
        return np.stack([o.tracer(o.old, l) for l in range(self.st.cfg.ntracers)])

    def forward(self, q=None):
        o = self.o.clone()
        if q is not None:
            for l in range(self.st.cfg.ntracers):
                o.set_tracer(o.old, l, q[l])
        for n in range(self.nsteps):
            self.wind(o, n)
            assert o.run(n) == 0
        return self.out(o)

    def tl(self, dq):
        o = self.o.clone()
        d = np.array(dq)
        for n in range(self.nsteps):
            self.wind(o, n)
            assert o.tlad_run(n, dq=d) == 0
        return d

    def ad(self, lam):
        states, o = [], self.o.clone()
        for n in range(self.nsteps):
            states.append(o.clone())
            self.wind(o, n)
            assert o.run(n) == 0
        a = np.array(lam)
        for n in reversed(range(self.nsteps)):
            s = states[n]
            self.wind(s, n)
            assert s.tlad_run(n, lam=a) == 0
        return a


@pytest.mark.parametrize("name", ["mono", "pd", "none"])
def test_ppm_jacobian(name):
    limiter = LIMITERS[name]
    rng = np.random.default_rng(11 + limiter)
    for _ in range(200):
        f = rng.uniform(0.1, 1, 5)
        J = oracle.ppm_jac(limiter, f)
        assert np.allclose(J @ f, oracle.ppm(f, limiter), rtol=0, atol=1e-14)
        e = 1e-7
        for t in range(5):
            fp, fm = f.copy(), f.copy()
            fp[t] += e
            fm[t] -= e
            fd = (np.array(oracle.ppm(fp, limiter)) - np.array(oracle.ppm(fm, limiter))) / (2 * e)
            assert np.abs(fd - J[:, t]).max() < 1e-7


@pytest.mark.parametrize("limiter,kind", [("mono", "smooth"), ("mono", "random"), ("none", "square")])
def test_tangent_euler_identity(limiter, kind):
    """J(q) q = F(q) over one advancing step (step 0 only accumulates)."""
    st = _setup(limiter)
    q0 = _fields(st, kind)
    w = Window(st, q0, 2)
    fq = w.forward()
    jq = w.tl(q0)
    assert np.abs(jq - fq).max() <= 1e-12 * np.abs(fq).max()
    if limiter == "none":
        o = w.o.clone()
        for n in range(2):
            w.wind(o, n)
            o.run(n)
        assert o.fixes > 0


@pytest.mark.parametrize("limiter,kind,nsteps,e", [("mono", "smooth", 4, 1e-7), ("mono", "random", 3, 1e-6),
                                                     ("pd", "random", 3, 1e-6), ("none", "square", 2, 1e-8)])
def test_tangent_central_differences(limiter, kind, nsteps, e):
    """e is below the distance to the nearest branch change along d (the
    square wave's flat regions put kinks close by; after the fixer has set
    cells to exactly zero, later steps sit on ties, where the step has no
    derivative, so that case takes one advancing step)."""
    st = _setup(limiter)
    q0 = _fields(st, kind)
    w = Window(st, q0, nsteps)
    d = _fields(st, "random", seed=7) - 0.5
    tl = w.tl(d)
    fd = (w.forward(q0 + e * d) - w.forward(q0 - e * d)) / (2 * e)
    err = np.abs(fd - tl)
    assert err.max() <= 1e-7 * np.abs(tl).max(), err.max()


@pytest.mark.parametrize("limiter,kind,nsteps", [("mono", "smooth", 5), ("mono", "random", 3),
                                                 ("none", "square", 4), ("pd", "random", 3)])
def test_adjoint_dot_product(limiter, kind, nsteps):
    st = _setup(limiter)
    q0 = _fields(st, kind)
    w = Window(st, q0, nsteps)
    d = _fields(st, "random", seed=21) - 0.5
    lam = _fields(st, "random", seed=22) - 0.5
    lhs = float(np.vdot(lam, w.tl(d)))
    rhs = float(np.vdot(w.ad(lam), d))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs), (lhs, rhs)
