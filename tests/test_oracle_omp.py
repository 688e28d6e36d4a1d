"""The all-cores CPU baseline (liboracle_omp.so: the oracle's independent row
loops on OpenMP threads) gives the same bits as the one-thread oracle."""
import numpy as np
import pytest

import oracle
from helpers import STATE, namelist, seeded_ic
from gmcore_amd.host import Setup


@pytest.mark.parametrize("opts", [dict(), dict(pv_scheme="tvd", use_div_damp=True, div_damp_order=4),
                                  dict(time_scheme="wrfrk3", coriolis_scheme=2)])
def test_openmp_oracle_is_bitwise_equal(opts):
    st = Setup(namelist(180, 91, 600.0, **opts))
    ic = seeded_ic(st, lmax=20, seed=0)
    runs = []
    for omp in (False, True):
        o = oracle.SWOracle(st.mesh_struct, st.cfg, omp=omp)
        for k, a in zip(("u_lon", "v_lat", "gz", "gzs"), ic):
            o.set(1, k, a)
        o.prepare(1)
        o.step(600.0, 3)
        runs.append({f: o.get(o.level, f) for f in STATE + ("u_f", "v_f", "gz_f")})
        runs[-1]["diag"] = o.diagnose()
        o.close()
    for f in runs[0]:
        assert np.array_equal(runs[0][f], runs[1][f]), f
