"""Python host of the GPU FFSL tracer advection (include/gmcore_adv.h).

``AdvSetup`` builds what adv_init / adv_allocate_tracers and lon_damp_init
set up for one cell-centred batch (src/dynamics/adv/adv_mod.F90:114-146,
334-406; src/dynamics/damp/lon_damp_mod.F90:43-68); ``Advection`` drives the
batch on the GPU the way adv_run is called (adv_mod.F90:152-312,
gmcore_adv_driver.F90:84-92).  The damping weights come from the Fortran host
library, the advection from libgmcore_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import GmcAdvConfig, dptr
from .host import GmcError, Mesh, load_host
from .namelist import Namelist


def lon_damp_weights(mesh: Mesh) -> np.ndarray:
    """wgt_full_lat of lon_damp_init, from the Fortran host."""
    L = load_host()
    L.gmch_lon_damp_weights.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    w = np.zeros(mesh.nlat)
    if L.gmch_lon_damp_weights(mesh.nlat, dptr(mesh.metrics), dptr(w)):
        raise ValueError("lon_damp_weights failed")
    return w


class AdvSetup:
    """Mesh, damping weights and gmc_adv_config of one advection batch."""

    def __init__(self, nl: Namelist, ntracers: int, lon_halo_width: int | None = None,
                 lon_damp_passes: int = 3):
        self.nl = nl
        self.mesh = Mesh(nl.num_lon, nl.num_lat, nl.radius, nl.omega)
        self.wgt = lon_damp_weights(self.mesh)
        hw = nl.lon_halo_width if lon_halo_width is None else lon_halo_width
        dt_dyn = nl.dt_dyn if nl.dt_dyn else nl.dt_adv
        self.cfg = GmcAdvConfig(radius=nl.radius, dt_adv=nl.dt_adv, nstep=int(nl.dt_adv / dt_dyn),
                                ntracers=ntracers, lon_halo_width=hw,
                                limiter=capi.LIMITERS[nl.limiter_type],
                                lon_damp_passes=lon_damp_passes,
                                flux_type=capi.FLUX_TYPES[nl.ffsl_flux_type],
                                lon_damp_wgt=dptr(self.wgt))
        self.mesh_struct = self.mesh.struct()


class Advection:
    """One FFSL advection batch on the GPU."""

    def __init__(self, setup: AdvSetup):
        self.setup = setup
        self.lib = capi.load_hip()
        self.nlon, self.nlat = setup.nl.num_lon, setup.nl.num_lat
        self.ntracers = setup.cfg.ntracers
        self.h = C.c_void_p()
        rc = self.lib.gmc_adv_create(C.byref(setup.mesh_struct), C.byref(setup.cfg), C.byref(self.h))
        if rc:
            raise GmcError(f"gmc_adv_create failed with code {rc}")

    def _check(self, rc, what):
        if rc:
            msg = self.lib.gmc_adv_last_error(self.h)
            raise GmcError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.lib.gmc_adv_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a, rows):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (rows, self.nlon):
            raise ValueError(f"shape {a.shape}, expected {(rows, self.nlon)}")
        return a

    def set_tracer(self, level: int, tracer: int, q):
        a = self._arr(q, self.nlat)
        self._check(self.lib.gmc_adv_upload_tracer(self.h, level, tracer, dptr(a), self.nlon), "upload")

    def tracer(self, level: int, tracer: int) -> np.ndarray:
        a = np.zeros((self.nlat, self.nlon))
        self._check(self.lib.gmc_adv_download_tracer(self.h, level, tracer, dptr(a), self.nlon), "download")
        return a

    def set_wind(self, u=None, v=None, mfx=None, mfy=None, m=None):
        n, h = self.nlat, self.nlat - 1
        arrs = [None if x is None else self._arr(x, r) for x, r in ((u, n), (v, h), (mfx, n), (mfy, h), (m, n))]
        ptrs = [None if a is None else dptr(a) for a in arrs]
        self._check(self.lib.gmc_adv_set_wind(self.h, *ptrs, self.nlon), "set_wind")

    def run(self, time_step: int):
        self._check(self.lib.gmc_adv_run(self.h, time_step), "adv_run")

    def init_dyn(self, model, level=None):
        """gmcore_run's accumulation before the time loop from a prepared
        ShallowWater context's state (gmc_adv_init_dyn)."""
        lv = model.level if level is None else level
        self._check(self.lib.gmc_adv_init_dyn(self.h, model.ctx, lv), "init_dyn")

    def run_dyn(self, model, time_step: int, level=None):
        """adv_run(block, new) on a ShallowWater context's state after its
        step (gmc_adv_run_dyn); the model's current level is that state."""
        lv = model.level if level is None else level
        self._check(self.lib.gmc_adv_run_dyn(self.h, model.ctx, lv, time_step), "run_dyn")

    def set_wind_basis(self, u, v):
        ub, vb = self._arr(u, self.nlat), self._arr(v, self.nlat - 1)
        self._check(self.lib.gmc_adv_set_wind_basis(self.h, dptr(ub), dptr(vb), self.nlon), "set_wind_basis")

    def drive(self, time_step0: int, scales):
        s = np.ascontiguousarray(scales, dtype=np.float64)
        self._check(self.lib.gmc_adv_drive(self.h, time_step0, dptr(s), len(s)), "drive")

    def _tlad_buf(self, a):
        shape = (self.ntracers, self.nlat, self.nlon)
        b = np.array(a, dtype=np.float64, order="C")
        if b.shape != shape:
            raise ValueError(f"expected shape {shape}, got {b.shape}")
        return b

    def tl_drive(self, time_step0: int, scales, dq) -> np.ndarray:
        """drive() carrying the tangent dq (ntracers, nlat, nlon) of the old
        level at entry; returns the tangent of the old level at exit."""
        s = np.ascontiguousarray(scales, dtype=np.float64)
        b = self._tlad_buf(dq)
        self._check(self.lib.gmc_adv_tl_drive(self.h, time_step0, dptr(s), len(s), dptr(b), self.nlon), "tl_drive")
        return b

    def ad_drive(self, time_step0: int, scales, lam, ckpt_every: int = 8) -> np.ndarray:
        """drive(), then the adjoint of the window: lam, the adjoint of the
        old level at exit, mapped to the adjoint of the old level at entry."""
        s = np.ascontiguousarray(scales, dtype=np.float64)
        b = self._tlad_buf(lam)
        self._check(self.lib.gmc_adv_ad_drive(self.h, time_step0, dptr(s), len(s), ckpt_every, dptr(b), self.nlon),
                    "ad_drive")
        return b

    def gradient(self, time_step0: int, scales, y, ckpt_every: int = 8, out=None):
        """(J, dJ/dq_0) of J = 1/2 sum (q_N - y)^2 over the window drive()
        would run; y has shape (ntracers, nlat, nlon).  out: a C-contiguous
        float64 array of that shape that receives the gradient (a
        minimisation reuses it from one call to the next); a new array when
        None."""
        s = np.ascontiguousarray(scales, dtype=np.float64)
        shape = (self.ntracers, self.nlat, self.nlon)
        yb = np.ascontiguousarray(y, dtype=np.float64)   # read only: no copy when already C-ordered float64
        if yb.shape != shape:
            raise ValueError(f"expected shape {shape}, got {yb.shape}")
        if out is None:
            grad = np.empty_like(yb)
        else:
            grad = out
            if grad.dtype != np.float64 or grad.shape != shape or not grad.flags.c_contiguous:
                raise ValueError(f"out: a C-contiguous float64 array of shape {shape}")
        cost = C.c_double()
        self._check(self.lib.gmc_adv_gradient(self.h, time_step0, dptr(s), len(s), ckpt_every, dptr(yb),
                                              dptr(grad), C.byref(cost), self.nlon), "gradient")
        return cost.value, grad

    @property
    def old(self) -> int:
        out = C.c_int32()
        self._check(self.lib.gmc_adv_old_level(self.h, C.byref(out)), "old_level")
        return out.value

    def mass(self, level=None) -> np.ndarray:
        out = np.zeros(self.ntracers)
        self._check(self.lib.gmc_adv_mass(self.h, self.old if level is None else level, dptr(out)), "mass")
        return out

    @property
    def fixes(self) -> int:
        out = C.c_int64()
        self._check(self.lib.gmc_adv_fix_count(self.h, C.byref(out)), "fix_count")
        return out.value

    def synchronize(self):
        self._check(self.lib.gmc_adv_synchronize(self.h), "synchronize")
