"""Shared set-up for the parity tests: one configuration run through the
oracle (CPU restatement) and through libgmcore_hip.so on the same inputs."""
import numpy as np

import oracle
from gmcore_amd import init_cond
from gmcore_amd.host import Setup, ShallowWater
from gmcore_amd.namelist import Namelist

STATE = ("u_lon", "v_lat", "gz")


def namelist(nlon=180, nlat=91, dt=600.0, **kw):
    nl = Namelist(num_lon=nlon, num_lat=nlat, dt_dyn=dt, time_scheme=kw.pop("time_scheme", "pc2"))
    for k, v in kw.items():
        setattr(nl, k, v)
    return nl


def seeded_ic(setup, lmax=20, seed=0, vmax=40.0):
    nl = setup.nl
    return init_cond.seeded_state(nl.num_lon, nl.num_lat, setup.mesh.metrics, lmax, seed, vmax=vmax,
                                  omega=nl.omega, radius=nl.radius)


def make_oracle(setup, ic):
    o = oracle.SWOracle(setup.mesh_struct, setup.cfg)
    for k, a in zip(("u_lon", "v_lat", "gz", "gzs"), ic):
        o.set(1, k, a)
    return o


def make_gpu(setup, ic):
    m = ShallowWater(setup)
    m.set_state(*ic)
    return m


def rel_err(a, b):
    """max |a - b| / max |b| (the per-field relative error of the parity bar)"""
    scale = np.abs(b).max()
    return float(np.abs(a - b).max() / (scale if scale > 0 else 1.0))
