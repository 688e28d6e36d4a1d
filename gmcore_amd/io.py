"""History (h0) and restart (r0) files through include/gmcore_io.h: the
netCDF classic files the reference's history_mod / restart_mod define
(src/utils/history_mod.F90:125-150, 398-444; src/utils/restart_mod.F90:
54-251), written and read by libgmcore_hip.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import GmcH0Record, dptr


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lon_degrees(nlon: int, dlon: float):
    """full_lon_deg, half_lon_deg (mesh_mod.F90:180-185, start_lon = 0)"""
    deg = 180.0 / np.pi
    full = np.arange(nlon) * dlon
    return full * deg, (full + 0.5 * dlon) * deg


def write_h0(path: str, fields: dict, coords: tuple, time_value: float, time_units: str, dt: float,
             totals: tuple, new_file: bool = True):
    """fields: u (nlat, nlon), v (nlat-1, nlon), z, pv (nlat-1, nlon), div, zs;
    coords: (lon, lat, ilon, ilat) in degrees; totals: (tm, te, tpe)."""
    f = {k: _d(v) for k, v in fields.items()}
    c = [_d(a) for a in coords]
    nlat, nlon = f["z"].shape
    units = time_units.encode()
    r = GmcH0Record(nlon, nlat, dptr(c[0]), dptr(c[1]), dptr(c[2]), dptr(c[3]), time_value, units, dt,
                    totals[0], totals[1], totals[2], dptr(f["u"]), dptr(f["v"]), dptr(f["z"]), dptr(f["pv"]),
                    dptr(f["div"]), dptr(f["zs"]))
    rc = capi.load_hip().gmc_h0_write(path.encode(), C.byref(r), int(new_file))
    if rc:
        raise OSError(f"gmc_h0_write({path}) failed ({rc})")


def write_r0(path: str, u, v, gz, gzs, coords: tuple, time_value: float, time_units: str, dt_dyn: float,
             restart_interval: str):
    a = [_d(x) for x in (u, v, gz, gzs)]
    c = [_d(x) for x in coords]
    nlat, nlon = a[2].shape
    rc = capi.load_hip().gmc_r0_write(path.encode(), nlon, nlat, *[dptr(x) for x in c], time_value,
                                      time_units.encode(), dt_dyn, restart_interval.encode(),
                                      *[dptr(x) for x in a])
    if rc:
        raise OSError(f"gmc_r0_write({path}) failed ({rc})")


def read_r0(path: str, nlon: int, nlat: int):
    """(u, v, gz, gzs, time_value) of record 0"""
    u, gz, gzs = (np.zeros((nlat, nlon)) for _ in range(3))
    v = np.zeros((nlat - 1, nlon))
    t = C.c_double()
    rc = capi.load_hip().gmc_r0_read(path.encode(), nlon, nlat, dptr(u), dptr(v), dptr(gz), dptr(gzs),
                                     C.byref(t))
    if rc:
        raise OSError(f"gmc_r0_read({path}) failed ({rc})")
    return u, v, gz, gzs, t.value
