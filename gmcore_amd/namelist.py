"""&gmcore_control keys used on the shallow-water / FFSL path.

Defaults follow src/utils/namelist_mod.F90:10-122.  Several of those
variables are real(8) but are initialised with default-real (single
precision) literals, e.g. ``filter_coef_b = 0.4`` (:92); the reference
therefore runs with the float32-rounded value.  ``_f32`` reproduces that.
Values read from a namelist file are parsed as double, as Fortran does.
"""
from __future__ import annotations

import dataclasses
import re

import numpy as np


def _f32(x: float) -> float:
    return float(np.float32(x))


# Earth constants, src/utils/const_mod.F90:55-66
PI = 3.141592653589793
EARTH_OMEGA = 2 * PI / 86400.0
EARTH_RADIUS = 6.37122e6
EARTH_G = 9.80616


@dataclasses.dataclass
class Namelist:
    num_lon: int = 0
    num_lat: int = 0
    num_lev: int = 1
    dt_dyn: float = 0.0
    dt_adv: float = 0.0
    time_scheme: str = "wrfrk3"          # namelist_mod.F90:83 (SW configs set 'pc2')
    pv_scheme: str = "upwind"            # :59
    pv_pole_stokes: bool = True          # :60
    upwind_order_pv: int = 3             # :61
    upwind_wgt_pv: float = 1.0           # :62
    coriolis_scheme: int = 1             # :65
    ke_scheme: int = 1                   # forced to 1 for one level, :228-235
    tangent_wgt_scheme: str = "classic"  # :45
    max_wave_speed: float = _f32(300)    # :89
    max_cfl: float = _f32(0.5)           # :90
    filter_coef_a: float = _f32(1.0)     # :91
    filter_coef_b: float = _f32(0.4)     # :92
    filter_coef_c: float = _f32(0.2)     # :93
    filter_coef_d: float = _f32(0.375)   # :94
    filter_coef_e: float = _f32(0.16)    # :95
    use_div_damp: bool = False           # :100
    div_damp_order: int = 2              # :101
    div_damp_coef2: float = 1.0 / 128.0  # :107
    div_damp_coef4: float = 0.01         # :108
    tvd_limiter_type: str = "van_leer"   # :70
    lon_halo_width: int = 2              # :38
    advection: bool = False
    limiter_type: str = "mono"           # :68
    ffsl_flux_type: str = "ppm"          # :69
    planet: str = "earth"

    @property
    def radius(self) -> float:
        return EARTH_RADIUS

    @property
    def omega(self) -> float:
        return EARTH_OMEGA

    @property
    def g(self) -> float:
        return EARTH_G


def parse_namelist(path: str) -> Namelist:
    """Read the keys above from a Fortran namelist file (&gmcore_control)."""
    text = open(path).read()
    body = re.search(r"&gmcore_control(.*?)/", text, re.S | re.I)
    nl = Namelist()
    if not body:
        return nl
    fields = {f.name: f for f in dataclasses.fields(Namelist)}
    for key, val in re.findall(r"(\w+)\s*=\s*([^,\n]+)", body.group(1)):
        key = key.lower()
        if key not in fields:
            continue
        val = val.strip().strip("'\"")
        cur = getattr(nl, key)
        if isinstance(cur, bool):
            setattr(nl, key, val.lower().strip(".") in ("t", "true"))
        elif isinstance(cur, int):
            setattr(nl, key, int(val))
        elif isinstance(cur, float):
            setattr(nl, key, float(val.lower().replace("d", "e")))
        else:
            setattr(nl, key, val)
    # parse_namelist, namelist_mod.F90:219-248
    if nl.num_lev == 1:
        nl.ke_scheme = 1
    if nl.dt_dyn == 0:
        nl.dt_dyn = nl.dt_adv
    if nl.dt_adv == 0:
        nl.dt_adv = nl.dt_dyn
    return nl
