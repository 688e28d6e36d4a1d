"""Set-up shared by the coupled dynamics + tracer tests: a 2-degree
shallow-water model (the seeded balanced state, pc2, dt = 600 s) and an FFSL
batch of cosine-bell tracers advanced with the model's winds every nstep
dynamics steps."""
import numpy as np

import oracle
from helpers import make_oracle, namelist, seeded_ic
from gmcore_amd import init_cond
from gmcore_amd.adv import AdvSetup
from gmcore_amd.host import Setup
from gmcore_amd.namelist import Namelist

NLON, NLAT, DT = 180, 91, 600.0


def setups(nstep, ntracers=3, flux="ppm", damp_passes=3):
    sw = Setup(namelist(NLON, NLAT, DT))
    nl = Namelist(num_lon=NLON, num_lat=NLAT, dt_dyn=DT, dt_adv=DT * nstep, ffsl_flux_type=flux)
    adv = AdvSetup(nl, ntracers, lon_halo_width=60, lon_damp_passes=damp_passes)
    return sw, adv


def initial(sw, ntracers=3):
    ic = seeded_ic(sw, lmax=20, seed=0)
    qs = init_cond.tracer_set(NLON, NLAT, sw.mesh.metrics, 3, ntracers)
    return ic, qs


def coupled_oracle(sw, adv, ic, qs):
    o = make_oracle(sw, ic)
    o.prepare(1)
    a = oracle.AdvOracle(adv.mesh_struct, adv.cfg)
    for l, q in enumerate(qs):
        a.set_tracer(1, l, q)
    c = oracle.CoupledOracle(o, a, DT)
    c.init()
    return c


def tracer_mass(q, m, sw):
    return float(np.sum(q * m * sw.mesh["area_cell"][:, None]))
