"""h0 and r0 files from a device context (gmc_history_h0, gmc_restart_write,
gmc_restart_read): the values the context holds, bit for bit, and a run
restarted from r0 that matches the oracle restarted from the same file
(restart_read then gmcore_run's set-up, restart_mod.F90:151-249,
gmcore_mod.F90:97-149)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from helpers import STATE, make_gpu, make_oracle, namelist, rel_err, seeded_ic
from gmcore_amd import io
from gmcore_amd.host import Setup

pytestmark = pytest.mark.gpu


def _model():
    st = Setup(namelist(72, 37, 1800.0))
    ic = seeded_ic(st, lmax=8, seed=2)
    g = make_gpu(st, ic)
    g.keep_derived(True)
    g.prepare(1)
    return st, ic, g


def test_h0_holds_the_context_fields(tmp_path):
    st, ic, g = _model()
    path = str(tmp_path / "swm.h0.nc")
    want = []
    for k in range(3):
        g.step(2)
        g.history_h0(path, 2.0 * (k + 1), "hours since 2000-01-01 00:00:00", new_file=k == 0)
        want.append((g.state(), g.download(g.level, "pv"), g.diagnose()))
    grav = st.nl.g
    with netcdf_file(path, "r", mmap=False) as nc:
        v = nc.variables
        assert v["u"].shape == (3, 37, 72) and v["pv"].shape == (3, 36, 72)
        for k, (s, pv, d) in enumerate(want):
            assert np.array_equal(v["u"][k], s["u_lon"])
            assert np.array_equal(v["v"][k], s["v_lat"])
            assert np.array_equal(v["z"][k], s["gz"] / grav)
            assert np.array_equal(v["pv"][k], pv)
            assert v["te"][k] == d[1] and v["tpe"][k] == d[2] and v["tm"][k] == np.float32(d[0])
        assert np.array_equal(v["zs"][:], (ic[3] / grav).astype(np.float32))
    g.close()


def test_restart_round_trip_and_continuation(tmp_path):
    st, ic, g = _model()
    g.step(3)
    path = str(tmp_path / "swm.r0.nc")
    g.restart_write(path, 3 * 1800.0, "seconds since 2000-01-01 00:00:00", "1 days")
    s3 = g.state()
    u, v, gz, gzs, t = io.read_r0(path, 72, 37)
    assert t == 3 * 1800.0
    assert np.array_equal(u, s3["u_lon"]) and np.array_equal(v, s3["v_lat"]) and np.array_equal(gz, s3["gz"])
    # a new context restarted from the file, and the oracle restarted from it
    g2 = make_gpu(st, (np.zeros_like(u), np.zeros_like(v), np.zeros_like(gz), np.zeros_like(gzs)))
    assert g2.restart_read(path, 1) == t
    g2.prepare(1)
    o = make_oracle(st, (u, v, gz, gzs))
    o.prepare(1)
    for n in range(2):
        g2.step(1)
        o.step(1800.0, 1)
        for f in STATE:
            assert rel_err(g2.download(g2.level, f), o.get(o.level, f)) <= 1e-12, (n, f)
    g.close()
    g2.close()
