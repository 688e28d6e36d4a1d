"""Hydrostatic column operators of the baroclinic dynamics on the GPU
(include/gmcore_baro.h): the Python mirror of operators_mod's calc_ph,
calc_t, calc_gz_lev, calc_dphsdt, calc_we_lev, calc_wedudlev_wedvdlev, the
vertical part of calc_grad_ptf, pgf_lin97_run (hydrostatic branch) and the
small_filter2 step of update_state.

Fields are numpy arrays [levels, rows, nlon]: levels nlev (full), nlev + 1
(half) or 1; rows nlat, or nlat - 1 for V-point fields (capi.BARO_FIELDS).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import BARO_FIELD, BARO_FIELDS, GmcThermo, GmcVcoord, dptr

# const_mod.F90:56-66 (Earth)
RD, CPD, P0 = 287.04, 1004.0, 1.0e5


class BaroError(RuntimeError):
    pass


class Baroclinic:
    """One block of nlev levels (gmc_baro_create).  hyai / hybi: the nlev + 1
    half-level coefficients of the hybrid coordinate; p0 its
    hybrid_coord_p0, ptop its local_ptop; Rd, cpd, tp0 the constants of
    temperature()."""

    def __init__(self, mesh, hyai, hybi, p0: float = P0, ptop: float = 0.0, Rd: float = RD, cpd: float = CPD,
                 tp0: float = P0):
        self.lib = capi.load_hip()
        self.mesh = mesh
        self.nlon, self.nlat = mesh.nlon, mesh.nlat
        self.hyai = np.ascontiguousarray(hyai, dtype=np.float64)
        self.hybi = np.ascontiguousarray(hybi, dtype=np.float64)
        if self.hyai.shape != self.hybi.shape or self.hyai.ndim != 1:
            raise ValueError("hyai and hybi must be 1-D of nlev + 1")
        self.nlev = self.hyai.size - 1
        vc = GmcVcoord(self.nlev, 0, p0, ptop, dptr(self.hyai), dptr(self.hybi))
        th = GmcThermo(Rd, cpd, tp0)
        self._mesh_struct = mesh.struct()
        h = C.c_void_p()
        rc = self.lib.gmc_baro_create(C.byref(self._mesh_struct), C.byref(vc), C.byref(th), C.byref(h))
        if rc:
            raise BaroError(f"gmc_baro_create failed ({rc})")
        self.h = h

    def _check(self, rc, what):
        if rc:
            msg = self.lib.gmc_baro_last_error(self.h)
            raise BaroError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def shape(self, name: str) -> tuple:
        lev, vrows = BARO_FIELDS[name]
        levels = 1 if lev == 0 else (self.nlev if lev == 1 else self.nlev + 1)
        return (levels, self.nlat - 1 if vrows else self.nlat, self.nlon)

    def set(self, name: str, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape(name))
        self._check(self.lib.gmc_baro_set(self.h, BARO_FIELD[name], dptr(a), self.nlon), f"set {name}")

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.shape(name))
        self._check(self.lib.gmc_baro_get(self.h, BARO_FIELD[name], dptr(out), self.nlon), f"get {name}")
        return out

    def calc_ph(self):
        self._check(self.lib.gmc_baro_calc_ph(self.h), "calc_ph")

    def calc_t(self):
        self._check(self.lib.gmc_baro_calc_t(self.h), "calc_t")

    def calc_gz_lev(self):
        self._check(self.lib.gmc_baro_calc_gz_lev(self.h), "calc_gz_lev")

    def calc_dphsdt(self):
        self._check(self.lib.gmc_baro_calc_dphsdt(self.h), "calc_dphsdt")

    def calc_we_lev(self):
        self._check(self.lib.gmc_baro_calc_we_lev(self.h), "calc_we_lev")

    def calc_wedudlev_wedvdlev(self):
        self._check(self.lib.gmc_baro_calc_wedudlev_wedvdlev(self.h), "calc_wedudlev_wedvdlev")

    def pgf_lin97(self):
        self._check(self.lib.gmc_baro_pgf_lin97(self.h), "pgf_lin97")

    def calc_ptf_lev(self, limiter: str = "mono", check: bool = True):
        """The vertical FFSL flux of pt (calc_grad_ptf's vertical part).  With
        check, wait for it and raise BaroError if a Courant number reached
        outside the column; otherwise the next synchronize() reports it."""
        self._check(self.lib.gmc_baro_calc_ptf_lev(self.h, capi.LIMITERS[limiter]), "calc_ptf_lev")
        if check:
            self.synchronize()

    def calc_ptf_hflx(self, dt: float, limiter: str = "mono", flux_type: str = "ppm", lon_halo_width: int = 1,
                      check: bool = True):
        """The horizontal part of calc_grad_ptf: ptf_lon, ptf_lat, dptfdlon,
        dptfdlat from pt and the mass batch's adv_* fields.  With check,
        wait and raise BaroError if a cflx exceeded lon_halo_width."""
        self._check(self.lib.gmc_baro_calc_ptf_hflx(self.h, capi.LIMITERS[limiter], capi.FLUX_TYPES[flux_type],
                                                    lon_halo_width, dt), "calc_ptf_hflx")
        if check:
            self.synchronize()

    def calc_grad_ptf(self, dt: float, limiter: str = "mono", flux_type: str = "ppm", lon_halo_width: int = 1,
                      check: bool = True):
        """calc_grad_ptf: the horizontal fluxes, then the vertical ones."""
        self._check(self.lib.gmc_baro_calc_grad_ptf(self.h, capi.LIMITERS[limiter], capi.FLUX_TYPES[flux_type],
                                                    lon_halo_width, dt), "calc_grad_ptf")
        if check:
            self.synchronize()

    def set_small_filter2(self, table):
        """Copy the small_filter2 table (a host.Filter of kind 2) to the
        device, once per run as filter_init builds it."""
        self._table = table.struct()
        self._check(self.lib.gmc_baro_set_small_filter2(self.h, C.byref(self._table)), "set_small_filter2")

    def small_filter2(self, table=None):
        """update_state's small_filter2 step on u_lon and v_lat with the table
        set by set_small_filter2 (or given here, which sets it first)."""
        if table is not None:
            self.set_small_filter2(table)
        self._check(self.lib.gmc_baro_small_filter2(self.h), "small_filter2")

    def synchronize(self):
        self._check(self.lib.gmc_baro_synchronize(self.h), "synchronize")

    def close(self):
        if getattr(self, "h", None):
            self.lib.gmc_baro_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
