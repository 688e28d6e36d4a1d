"""GPU tangent linear and adjoint of FFSL advection (gmc_adv_tl_drive,
gmc_adv_ad_drive) against the oracle's (oracle/adv.c, gmo_adv_tlad_run),
which tests/test_oracle_adv_tlad.py pins by Euler's identity, central
differences and the dot-product test.

Bars (FP64): the tangent and the adjoint within 1e-12 of the oracle's
(relative to the field's largest value; the GPU forms the same branch
choices from bit-identical base states, but sums the linear terms in its own
order, and the adjoint's scatters add with atomics); the dot-product test
<l, A d> = <A^T l, d> to 1e-12; Euler's identity A(q) q = F(q) to 1e-12;
the object's state after either call bit-identical to gmc_adv_drive's.  At
0.5 degrees with 8 tracers (the oracle's adjoint would take minutes there)
the size-independent properties are checked: dot product, Euler's identity
and checkpoint-interval independence.
"""
import numpy as np
import pytest

import oracle
from gmcore_amd import init_cond
from gmcore_amd.adv import Advection, AdvSetup
from gmcore_amd.host import GmcError
from gmcore_amd.namelist import Namelist

pytestmark = pytest.mark.gpu


def _setup(nlon, nlat, dt, ntr, hw, limiter, nstep=1):
    nl = Namelist(num_lon=nlon, num_lat=nlat, dt_adv=dt, dt_dyn=dt / nstep, limiter_type=limiter)
    return AdvSetup(nl, ntr, lon_halo_width=hw, lon_damp_passes=3)


def _fields(st, kind, seed):
    nlat, nlon, ntr = st.nl.num_lat, st.nl.num_lon, st.cfg.ntracers
    rng = np.random.default_rng(seed)
    if kind == "bells":
        return np.stack(init_cond.tracer_set(nlon, nlat, st.mesh.metrics, seed, ntr))
    if kind == "square":
        q = np.zeros((ntr, nlat, nlon))
        q[:, nlat // 3:2 * nlat // 3, nlon // 4:nlon // 2] = 1.0
        return q + 1e-3 * rng.uniform(0, 1, q.shape)
    return rng.uniform(0, 1, (ntr, nlat, nlon))


class Case:
    def __init__(self, nlon, nlat, dt, ntr, hw, limiter, kind, nsteps, lmax=6, vmax=40.0, nstep=1):
        self.st = _setup(nlon, nlat, dt, ntr, hw, limiter, nstep)
        self.nlat, self.nlon, self.ntr, self.nsteps = nlat, nlon, ntr, nsteps
        self.q0 = _fields(self.st, kind, 3)
        self.ub, self.vb = init_cond.deform_wind_basis(nlon, nlat, self.st.mesh.metrics, lmax, 7, vmax)
        self.sc = init_cond.wind_scales(nsteps, dt)
        self.d = _fields(self.st, "random", 21) - 0.5
        self.lam = _fields(self.st, "random", 22) - 0.5

    def gpu(self):
        g = Advection(self.st)
        for l in range(self.ntr):
            g.set_tracer(1, l, self.q0[l])
            g.set_tracer(2, l, self.q0[l])
        g.set_wind_basis(self.ub, self.vb)
        return g

    def final(self, g):
        return np.stack([g.tracer(g.old, l) for l in range(self.ntr)])

    def oracle(self):
        o = oracle.AdvOracle(self.st.mesh_struct, self.st.cfg)
        for l in range(self.ntr):
            o.set_tracer(1, l, self.q0[l])
            o.set_tracer(2, l, self.q0[l])
        return o

    def owind(self, o, n):
        u, v = self.ub * self.sc[n], self.vb * self.sc[n]
        o.set_wind(u, v, u, v, np.ones((self.nlat, self.nlon)))

    def oracle_tl(self, d):
        o, d = self.oracle(), np.array(d)
        for n in range(self.nsteps):
            self.owind(o, n)
            assert o.tlad_run(n, dq=d) == 0
        return d, np.stack([o.tracer(o.old, l) for l in range(self.ntr)])

    def oracle_ad(self, lam):
        o, states = self.oracle(), []
        for n in range(self.nsteps):
            states.append(o.clone())
            self.owind(o, n)
            assert o.run(n) == 0
        a = np.array(lam)
        for n in reversed(range(self.nsteps)):
            self.owind(states[n], n)
            assert states[n].tlad_run(n, lam=a) == 0
        return a


SMALL = {
    "mono-smooth": (72, 37, 3600.0, 2, 20, "mono", "bells", 5),
    "none-square-fixer": (72, 37, 3600.0, 2, 20, "none", "square", 4),
    "pd-random": (72, 37, 3600.0, 2, 20, "pd", "random", 3),
    "mono-180x91-3tr": (180, 91, 1800.0, 3, 60, "mono", "bells", 6),
}


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("name", list(SMALL))
def test_tangent_matches_oracle(name):
    c = Case(*SMALL[name])
    ref, qref = c.oracle_tl(c.d)
    g = c.gpu()
    tl = g.tl_drive(0, c.sc, c.d)
    assert np.array_equal(c.final(g), qref), "base trajectory differs from the oracle's"
    assert _rel(tl, ref) <= 1e-12, _rel(tl, ref)
    if name.startswith("none"):
        assert g.fixes > 0


@pytest.mark.parametrize("name", list(SMALL))
def test_adjoint_matches_oracle(name):
    c = Case(*SMALL[name])
    ref = c.oracle_ad(c.lam)
    g = c.gpu()
    ad = g.ad_drive(0, c.sc, c.lam, ckpt_every=2)
    g2 = c.gpu()
    g2.drive(0, c.sc)
    assert np.array_equal(c.final(g), c.final(g2)), "object state after the adjoint differs from drive's"
    assert g.fixes == g2.fixes
    assert _rel(ad, ref) <= 1e-12, _rel(ad, ref)


BIG = (720, 361, 900.0, 8, 104, "mono", "bells", 4, 10, 50.0)


@pytest.mark.parametrize("case", [SMALL["none-square-fixer"], BIG], ids=["72x37-fixer", "720x361-8tr"])
def test_dot_product_and_euler(case):
    c = Case(*case)
    tl = c.gpu().tl_drive(0, c.sc, c.d)
    ad = c.gpu().ad_drive(0, c.sc, c.lam, ckpt_every=3)
    lhs, rhs = float(np.vdot(c.lam, tl)), float(np.vdot(ad, c.d))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    g = c.gpu()
    jq = g.tl_drive(0, c.sc, c.q0)
    fq = c.final(g)
    assert _rel(jq, fq) <= 1e-12, _rel(jq, fq)


def test_multi_step_batch_adjoint():
    """A batch of nstep = 2 advances at every second time step: the adjoint
    tapes only the advancing steps, whichever checkpoint interval, and still
    matches the oracle and the tangent's dot product."""
    c = Case(72, 37, 3600.0, 2, 20, "mono", "bells", 7, nstep=2)
    assert c.st.cfg.nstep == 2
    ref = c.oracle_ad(c.lam)
    outs = [c.gpu().ad_drive(0, c.sc, c.lam, ckpt_every=k) for k in (1, 2, 3, 100)]
    for o in outs:
        assert _rel(o, ref) <= 1e-12, _rel(o, ref)
    tl = c.gpu().tl_drive(0, c.sc, c.d)
    lhs, rhs = float(np.vdot(c.lam, tl)), float(np.vdot(outs[0], c.d))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


def test_checkpoint_interval_independent():
    c = Case(*SMALL["mono-180x91-3tr"])
    outs = [c.gpu().ad_drive(0, c.sc, c.lam, ckpt_every=k) for k in (1, 4, 100)]
    for o in outs[1:]:
        assert _rel(o, outs[0]) <= 1e-14


@pytest.mark.parametrize("case", [(72, 37, 3600.0, 2, 20, "mono", "random", 4), BIG], ids=["72x37", "720x361-8tr"])
def test_gradient_of_final_misfit(case):
    """gmc_adv_gradient: J = 1/2 |q_N - y|^2 and dJ/dq_0.  J against the
    forward's end state; the gradient against gmc_adv_ad_drive forced by
    q_N - y, and (on the tie-free random base) against central differences
    of J along a direction: J is quadratic on each linear piece of the step,
    so the quotient is exact there up to rounding.  (The cosine bells' flat
    background puts the limiters on ties, where the step has one-sided
    derivatives only.)"""
    c = Case(*case)
    y = np.stack(init_cond.tracer_set(c.nlon, c.nlat, c.st.mesh.metrics, 9, c.ntr))
    g = c.gpu()
    J, grad = g.gradient(0, c.sc, y, ckpt_every=3)
    g2 = c.gpu()
    g2.drive(0, c.sc)
    qN = c.final(g2)
    assert np.array_equal(c.final(g), qN)
    jref = 0.5 * float(np.sum((qN - y) ** 2))
    assert abs(J - jref) <= 1e-13 * jref
    ad = c.gpu().ad_drive(0, c.sc, qN - y, ckpt_every=2)
    assert _rel(grad, ad) <= 1e-13, _rel(grad, ad)
    if c.nlon > 72:
        return
    e = 1e-6

    def cost(q0):
        h = Advection(c.st)
        for l in range(c.ntr):
            h.set_tracer(1, l, q0[l])
            h.set_tracer(2, l, c.q0[l])
        h.set_wind_basis(c.ub, c.vb)
        h.drive(0, c.sc)
        return 0.5 * float(np.sum((c.final(h) - y) ** 2))

    fd = (cost(c.q0 + e * c.d) - cost(c.q0 - e * c.d)) / (2 * e)
    gd = float(np.vdot(grad, c.d))
    assert abs(fd - gd) <= 1e-6 * abs(gd), (fd, gd)


def test_window_without_advancing_step_and_bad_arguments():
    """Time step 0 only accumulates the wind (adv_mod.F90:166): a one-step
    window from step 0, and an empty window, leave the tangent and the
    adjoint as they are.  ckpt_every < 1 and a wrongly shaped array are
    rejected."""
    c = Case(*SMALL["mono-smooth"])
    assert np.array_equal(c.gpu().tl_drive(0, c.sc[:1], c.d), c.d)
    assert np.array_equal(c.gpu().ad_drive(0, c.sc[:1], c.lam, ckpt_every=4), c.lam)
    assert np.array_equal(c.gpu().ad_drive(0, np.zeros(0), c.lam, ckpt_every=4), c.lam)
    with pytest.raises(GmcError):
        c.gpu().ad_drive(0, c.sc, c.lam, ckpt_every=0)
    with pytest.raises(ValueError):
        c.gpu().tl_drive(0, c.sc, c.d[:, :-1])


def test_second_window_matches_oracle():
    """A window that starts after 3 plain steps, mid-way through the wind
    accumulation's cadence, against the oracle's tangent and adjoint of the
    same steps."""
    c = Case(72, 37, 3600.0, 2, 20, "mono", "random", 8)
    n0, n1 = 3, 8
    o = c.oracle()
    for n in range(n0):
        c.owind(o, n)
        assert o.run(n) == 0
    ot, d = o.clone(), np.array(c.d)
    for n in range(n0, n1):
        c.owind(ot, n)
        assert ot.tlad_run(n, dq=d) == 0
    states, oa = [], o.clone()
    for n in range(n0, n1):
        states.append(oa.clone())
        c.owind(oa, n)
        assert oa.run(n) == 0
    lam = np.array(c.lam)
    for k in reversed(range(n1 - n0)):
        c.owind(states[k], n0 + k)
        assert states[k].tlad_run(n0 + k, lam=lam) == 0
    g = c.gpu()
    g.drive(0, c.sc[:n0])
    tl = g.tl_drive(n0, c.sc[n0:n1], c.d)
    assert _rel(tl, d) <= 1e-12, _rel(tl, d)
    g = c.gpu()
    g.drive(0, c.sc[:n0])
    ad = g.ad_drive(n0, c.sc[n0:n1], c.lam, ckpt_every=2)
    assert _rel(ad, lam) <= 1e-12, _rel(ad, lam)
