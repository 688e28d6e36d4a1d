"""Python host of the GPU dynamics path.

The mesh metrics and filter tables come from the Fortran host library
(libgmcore_host.so, gmcore_amd/fortran), the dynamics from libgmcore_hip.so
through its C ABI.  ``ShallowWater`` drives one block the way gmcore_run does
(src/dynamics/gmcore_mod.F90:97-149): prepare, then time_integrate + swap.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi
from .capi import (GMC_MESH_NFIELDS, FIELD, GmcConfig, GmcFilter, GmcMesh, dptr, field_rows, iptr)
from .namelist import Namelist

_host = None


def load_host() -> C.CDLL:
    global _host
    if _host is None:
        if not os.path.exists(capi.LIB_HOST):
            raise RuntimeError(f"{capi.LIB_HOST} is not built; run __graft_entry__.build() first")
        L = C.CDLL(capi.LIB_HOST)
        P = C.POINTER
        L.gmch_mesh_metrics.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, P(C.c_double),
                                        P(C.c_double)]
        L.gmch_filter.argtypes = [C.c_int, P(C.c_double), C.c_double, P(C.c_double), C.c_int,
                                  P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_double),
                                  P(C.c_double)]
        _host = L
    return _host


class Mesh:
    """Packed metric table of mesh_type (mesh_mod.F90:134-447) for nlon x nlat."""

    def __init__(self, nlon: int, nlat: int, radius: float, omega: float):
        self.nlon, self.nlat = nlon, nlat
        self.metrics = np.zeros((GMC_MESH_NFIELDS, nlat))
        dlon = C.c_double()
        rc = load_host().gmch_mesh_metrics(nlon, nlat, radius, omega, dptr(self.metrics),
                                           C.byref(dlon))
        if rc:
            raise ValueError(f"mesh setup failed for {nlon}x{nlat}")
        self.dlon = dlon.value

    def __getitem__(self, name):
        return self.metrics[capi.MESH[name]]

    def struct(self) -> GmcMesh:
        return GmcMesh(self.nlon, self.nlat, self.dlon, dptr(self.metrics))


class Filter:
    """filter_type tables (filter_types_mod.F90:49-147); kind 0 big, 1 small_filter1,
    2 small_filter2."""

    def __init__(self, mesh: Mesh, dt: float, nl: Namelist, kind: int):
        nlat = mesh.nlat
        coefs = np.array([nl.max_wave_speed, nl.max_cfl, nl.filter_coef_a, nl.filter_coef_b,
                          nl.filter_coef_c, nl.filter_coef_d, nl.filter_coef_e])
        self.ngrid_lon = np.zeros(nlat, np.int32)
        self.ngrid_lat = np.zeros(nlat - 1, np.int32)
        mx = np.zeros(1, np.int32)
        L = load_host()
        args = (nlat, dptr(mesh.metrics), dt, dptr(coefs), kind, iptr(self.ngrid_lon),
                iptr(self.ngrid_lat), iptr(mx))
        dummy = np.zeros(1)
        if L.gmch_filter(*args, dptr(dummy), dptr(dummy)):
            raise ValueError("filter setup failed")
        self.max_taps = int(mx[0])
        self.wgt_lon = np.zeros((nlat, self.max_taps))
        self.wgt_lat = np.zeros((nlat - 1, self.max_taps))
        if L.gmch_filter(*args, dptr(self.wgt_lon), dptr(self.wgt_lat)):
            raise ValueError("filter setup failed")

    def struct(self) -> GmcFilter:
        return GmcFilter(self.max_taps, iptr(self.ngrid_lon), dptr(self.wgt_lon),
                         iptr(self.ngrid_lat), dptr(self.wgt_lat))


def make_config(nl: Namelist, big: Filter, small: Filter) -> GmcConfig:
    ts = {"pc2": capi.GMC_TIME_PC2, "wrfrk3": capi.GMC_TIME_WRFRK3, "euler": capi.GMC_TIME_EULER}
    if nl.pv_scheme not in capi.PV_SCHEMES:
        raise ValueError(f"pv_scheme {nl.pv_scheme!r} is not built (midpoint, upwind, tvd)")
    if nl.tvd_limiter_type not in capi.TVD_LIMITERS:
        raise ValueError(f"Invalid tvd_limiter_type {nl.tvd_limiter_type!r}")
    return GmcConfig(
        g=nl.g, time_scheme=ts.get(nl.time_scheme, capi.GMC_TIME_PC2),
        pv_scheme=capi.PV_SCHEMES[nl.pv_scheme],
        upwind_order_pv=nl.upwind_order_pv, upwind_wgt_pv=nl.upwind_wgt_pv,
        pv_pole_stokes=int(nl.pv_pole_stokes), coriolis_scheme=nl.coriolis_scheme,
        use_div_damp=int(nl.use_div_damp), div_damp_order=nl.div_damp_order,
        div_damp_coef2=nl.div_damp_coef2, big_filter=big.struct(), small_filter1=small.struct(),
        div_damp_coef4=nl.div_damp_coef4, dt_dyn=nl.dt_dyn,
        tvd_limiter=capi.TVD_LIMITERS[nl.tvd_limiter_type], reserved=0)


class Setup:
    """Mesh, filters and gmc_config of one configuration (what gmcore_init builds)."""

    def __init__(self, nl: Namelist):
        self.nl = nl
        self.mesh = Mesh(nl.num_lon, nl.num_lat, nl.radius, nl.omega)
        self.big = Filter(self.mesh, nl.dt_dyn, nl, 0)
        self.small = Filter(self.mesh, nl.dt_dyn, nl, 1)
        self.cfg = make_config(nl, self.big, self.small)
        self.mesh_struct = self.mesh.struct()


class GmcError(RuntimeError):
    pass


class ShallowWater:
    """One shallow-water block on the GPU (libgmcore_hip.so)."""

    def __init__(self, setup: Setup):
        self.setup = setup
        self.lib = capi.load_hip()
        self.nlon, self.nlat = setup.nl.num_lon, setup.nl.num_lat
        self.ctx = C.c_void_p()
        rc = self.lib.gmc_create(C.byref(setup.mesh_struct), C.byref(setup.cfg), C.byref(self.ctx))
        if rc:
            raise GmcError(f"gmc_create failed with code {rc}")

    def _check(self, rc, what):
        if rc:
            msg = self.lib.gmc_last_error(self.ctx)
            raise GmcError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self.ctx:
            self.lib.gmc_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, level: int, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != (field_rows(name, self.nlat), self.nlon):
            raise ValueError(f"{name}: shape {a.shape}")
        self._check(self.lib.gmc_upload(self.ctx, level, FIELD[name], dptr(a), self.nlon), "upload")

    def download(self, level: int, name: str) -> np.ndarray:
        a = np.zeros((field_rows(name, self.nlat), self.nlon))
        self._check(self.lib.gmc_download(self.ctx, level, FIELD[name], dptr(a), self.nlon), "download")
        return a

    def set_state(self, u, v, gz, gzs, level=1):
        self.upload(level, "u_lon", u)
        self.upload(level, "v_lat", v)
        self.upload(level, "gz", gz)
        self.upload(level, "gzs", gzs)

    def prepare(self, level=1):
        self._check(self.lib.gmc_prepare(self.ctx, level), "prepare")

    def step(self, nsteps=1, dt=None):
        self._check(self.lib.gmc_step(self.ctx, self.setup.nl.dt_dyn if dt is None else dt, nsteps),
                    "step")

    def time_integrate(self, old, new, dt=None):
        self._check(self.lib.gmc_time_integrate(self.ctx, old, new,
                                                self.setup.nl.dt_dyn if dt is None else dt),
                    "time_integrate")

    def space_operators(self, old, star, new, t1, t2, dt, pass_):
        self._check(self.lib.gmc_space_operators(self.ctx, old, star, new, t1, t2, dt, pass_),
                    "space_operators")

    def update_state(self, tend, old, new, dt, what):
        self._check(self.lib.gmc_update_state(self.ctx, tend, old, new, dt, what), "update_state")

    @property
    def level(self) -> int:
        out = C.c_int32()
        self._check(self.lib.gmc_current_level(self.ctx, C.byref(out)), "current_level")
        return out.value

    def state(self, level=None):
        level = self.level if level is None else level
        return {k: self.download(level, k) for k in ("u_lon", "v_lat", "gz")}

    def diagnose(self, level=None):
        out = np.zeros(4)
        self._check(self.lib.gmc_diagnose(self.ctx, self.level if level is None else level, dptr(out)),
                    "diagnose")
        return out

    def synchronize(self):
        self._check(self.lib.gmc_synchronize(self.ctx), "synchronize")

    def history_h0(self, path: str, time_value: float, time_units: str, new_file: bool, level=None):
        """Write (or append) one h0 record of the level (gmc_history_h0)."""
        lv = self.level if level is None else level
        self._check(self.lib.gmc_history_h0(self.ctx, path.encode(), lv, time_value, time_units.encode(),
                                            self.setup.nl.dt_dyn, int(new_file)), "history_h0")

    def restart_write(self, path: str, time_value: float, time_units: str, restart_interval: str, level=None):
        lv = self.level if level is None else level
        self._check(self.lib.gmc_restart_write(self.ctx, path.encode(), lv, time_value, time_units.encode(),
                                               self.setup.nl.dt_dyn, restart_interval.encode()), "restart_write")

    def restart_read(self, path: str, level: int = 1) -> float:
        """u, v, gz, gzs of an r0 file into the level; gmc_prepare must follow."""
        t = C.c_double()
        self._check(self.lib.gmc_restart_read(self.ctx, path.encode(), level, C.byref(t)), "restart_read")
        return t.value

    def keep_derived(self, on: bool = True):
        """Store each stage's forward-pass fields in the star level (gmc_keep_derived)."""
        self._check(self.lib.gmc_keep_derived(self.ctx, int(on)), "keep_derived")

    def filter_plan(self) -> dict:
        """Shape of the fused step's filter work list (gmc_filter_plan)."""
        out = (C.c_int32 * 4)()
        self._check(self.lib.gmc_filter_plan(self.ctx, out), "filter_plan")
        return {"items": out[0], "rows_split": out[1], "max_chunks": out[2], "widest": out[3]}


class Ensemble:
    """Batched members (include/gmcore_ens.h): one graph of three launches per
    stage steps every member, member k with its own dt."""

    def __init__(self, members):
        self.members = list(members)
        self.lib = capi.load_hip()
        arr = (C.c_void_p * len(self.members))(*[m.ctx.value for m in self.members])
        self.h = C.c_void_p()
        rc = self.lib.gmc_ens_create(arr, len(self.members), C.byref(self.h))
        if rc:
            raise GmcError(f"gmc_ens_create failed with code {rc}")
        self.dt = np.array([m.setup.nl.dt_dyn for m in self.members])

    def _check(self, rc, what):
        if rc:
            msg = self.lib.gmc_ens_last_error(self.h)
            raise GmcError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.lib.gmc_ens_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps=1):
        self._check(self.lib.gmc_ens_step(self.h, dptr(self.dt), nsteps), "ens_step")

    def set_shape(self, shape: int):
        self._check(self.lib.gmc_ens_shape(self.h, shape), "ens_shape")

    def autotune(self, trial_steps=3):
        """(chosen shape, mean step ms of every shape)"""
        n = 32
        t = np.zeros(n)
        choice, ncand = C.c_int32(), C.c_int32()
        self._check(self.lib.gmc_ens_autotune(self.h, dptr(self.dt), trial_steps, C.byref(choice), dptr(t), n,
                                              C.byref(ncand)), "ens_autotune")
        return choice.value, t[:ncand.value]

    @staticmethod
    def shape_name(k: int) -> str:
        return capi.load_hip().gmc_ens_shape_name(k).decode()


TLAD_VARS = {"u": 0, "v": 1, "gz": 2, "u_f": 3, "v_f": 4}


class TLAD:
    """Tangent-linear and adjoint variables of one ShallowWater context
    (include/gmcore_tlad.h)."""

    def __init__(self, model: ShallowWater):
        self.model = model
        self.lib = model.lib
        self.h = C.c_void_p()
        rc = self.lib.gmc_tlad_create(model.ctx, C.byref(self.h))
        if rc:
            raise GmcError(f"gmc_tlad_create failed with code {rc}")

    def _check(self, rc, what):
        if rc:
            msg = self.lib.gmc_tlad_last_error(self.h)
            raise GmcError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.lib.gmc_tlad_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _rows(self, var):
        nlat = self.model.nlat
        return nlat - 1 if var in ("v", "v_f") else nlat

    def set(self, level, var, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == (self._rows(var), self.model.nlon)
        self._check(self.lib.gmc_tlad_upload(self.h, level, TLAD_VARS[var], dptr(a), self.model.nlon), "upload")

    def get(self, level, var):
        a = np.zeros((self._rows(var), self.model.nlon))
        self._check(self.lib.gmc_tlad_download(self.h, level, TLAD_VARS[var], dptr(a), self.model.nlon), "download")
        return a

    def zero(self):
        self._check(self.lib.gmc_tlad_zero(self.h), "zero")

    def tl_step(self, nsteps=1, dt=None):
        self._check(self.lib.gmc_tl_step(self.h, self.model.setup.nl.dt_dyn if dt is None else dt, nsteps), "tl_step")

    def ad_step(self, nsteps=1, dt=None):
        self._check(self.lib.gmc_ad_step(self.h, self.model.setup.nl.dt_dyn if dt is None else dt, nsteps), "ad_step")

    def gradient(self, nsteps, ckpt_every, obs_cell, obs_value, sigma_o, xb, sigma_b, dt=None, out=None):
        """(J, (dJ/du0, dJ/dv0, dJ/dgz0)); xb = (u_b, v_b, gz_b).  out: three
        C-contiguous float64 arrays shaped like xb that receive the gradient
        (a minimisation reuses them from one call to the next); new arrays
        when None."""
        m = self.model
        cell = np.ascontiguousarray(obs_cell, dtype=np.int64)
        y = np.ascontiguousarray(obs_value, dtype=np.float64)
        ub, vb, gb = (np.ascontiguousarray(a, dtype=np.float64) for a in xb)
        if out is None:
            gu, gv, gg = np.empty_like(ub), np.empty_like(vb), np.empty_like(gb)   # every element is written
        else:
            gu, gv, gg = out
            for o, a in zip(out, (ub, vb, gb)):
                if o.dtype != np.float64 or o.shape != a.shape or not o.flags.c_contiguous:
                    raise ValueError("out: C-contiguous float64 arrays shaped like xb")
        J = C.c_double()
        self._check(self.lib.gmc_gradient(self.h, m.setup.nl.dt_dyn if dt is None else dt, nsteps, ckpt_every,
                                          len(cell), cell.ctypes.data_as(C.POINTER(C.c_int64)), dptr(y), sigma_o,
                                          dptr(ub), dptr(vb), dptr(gb), sigma_b, m.nlon, C.byref(J), dptr(gu),
                                          dptr(gv), dptr(gg)), "gradient")
        return J.value, (gu, gv, gg)
