"""The Fortran hosts on the GPU path, run as programs: gmcore_swm_gpu (shaped
like gmcore_swm_driver.F90) and gmcore_adv_gpu (shaped like
gmcore_adv_driver.F90) read a &gmcore_control namelist and raw float64
inputs, step through libgmcore_hip.so via ISO_C_BINDING, and write their
final state; the result must equal the oracle's on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from gmcore_amd import init_cond
from gmcore_amd.adv import AdvSetup
from gmcore_amd.capi import LIB_HIP
from gmcore_amd.host import Setup
from gmcore_amd.namelist import Namelist

from helpers import make_oracle, namelist, seeded_ic

pytestmark = pytest.mark.gpu
LIB = os.path.dirname(LIB_HIP)


def _run(exe, args, tmp_path):
    print(f"running {exe} {' '.join(map(str, args))}", flush=True)
    r = subprocess.run([os.path.join(LIB, exe), *map(str, args)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_swm_driver_matches_oracle(tmp_path):
    nlon, nlat, dt, steps = 72, 37, 1800.0, 4
    nml = tmp_path / "swm.nml"
    nml.write_text(f"&gmcore_control\n num_lon = {nlon}\n num_lat = {nlat}\n dt_dyn = {dt}\n"
                   f" run_steps = {steps}\n time_scheme = 'pc2'\n/\n")
    st = Setup(namelist(nlon, nlat, dt))
    ic = seeded_ic(st, lmax=8, seed=13)
    with open(tmp_path / "ic.bin", "wb") as f:
        for a in ic:
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
    out = _run("gmcore_swm_gpu", [nml, tmp_path / "ic.bin", tmp_path / "out.bin"], tmp_path)
    assert "tm =" in out
    raw = np.fromfile(tmp_path / "out.bin", dtype="<f8")
    n1, n2 = nlat * nlon, (nlat - 1) * nlon
    u, v, gz = raw[:n1].reshape(nlat, nlon), raw[n1:n1 + n2].reshape(nlat - 1, nlon), \
        raw[n1 + n2:2 * n1 + n2].reshape(nlat, nlon)
    o = make_oracle(st, ic)
    o.prepare(1)
    o.step(dt, steps)
    for name, a in (("u_lon", u), ("v_lat", v), ("gz", gz)):
        b = o.get(o.level, name)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), name


def test_adv_driver_matches_oracle(tmp_path):
    nlon, nlat, dt, steps, ntr = 72, 37, 3600.0, 6, 3
    nml = tmp_path / "adv.nml"
    nml.write_text(f"&gmcore_control\n num_lon = {nlon}\n num_lat = {nlat}\n dt_adv = {dt}\n"
                   f" run_steps = {steps}\n lon_halo_width = 20\n limiter_type = 'mono'\n/\n")
    nl = Namelist(num_lon=nlon, num_lat=nlat, dt_adv=dt, dt_dyn=dt)
    st = AdvSetup(nl, ntr, lon_halo_width=20)
    m = st.mesh.metrics
    ub, vb = init_cond.deform_wind_basis(nlon, nlat, m, 6, 5, 40.0)
    qs = init_cond.tracer_set(nlon, nlat, m, 6, ntr)
    with open(tmp_path / "in.bin", "wb") as f:
        for a in (ub, vb, *qs):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
    _run("gmcore_adv_gpu", [nml, tmp_path / "in.bin", tmp_path / "out.bin", ntr], tmp_path)
    got = np.fromfile(tmp_path / "out.bin", dtype="<f8").reshape(ntr, nlat, nlon)
    o = oracle.AdvOracle(st.mesh_struct, st.cfg)
    for l, q in enumerate(qs):
        o.set_tracer(1, l, q)
    sc = init_cond.wind_scales(steps, dt)
    ub0 = ub.copy()
    ub0[[0, -1]] = 0
    ones = np.ones((nlat, nlon))
    for n in range(steps):
        u, v = ub0 * sc[n], vb * sc[n]
        o.set_wind(u, v, u, v, ones)
        assert o.run(n) == 0
    for l in range(ntr):
        assert np.array_equal(got[l], o.tracer(o.old, l)), l
