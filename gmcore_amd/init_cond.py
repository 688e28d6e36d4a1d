"""Seeded shallow-water initial states (synthetic data, no files).

Recipe (the benchmark configurations): a streamfunction psi on the vertices
from random spherical harmonics l <= lmax with coefficient variance l**-3,
winds from discrete differences of psi (so the discrete divergence of
(u, v) is exactly zero, pole rows included), scaled to a given peak wind;
geopotential from the linear balance lap(Phi) = div(f grad psi) solved on
the C grid (FFT in longitude, tridiagonal in latitude), mean 5.0e4 m2/s2;
flat topography.  Also the analytic cases the reference uses as tests:
steady geostrophic flow (src/tests/swm/steady_geostrophic_flow_test_mod.F90)
and the Rossby-Haurwitz wave (rossby_haurwitz_wave_test_mod.F90).
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import solve_banded

from .capi import MESH


def legendre_table(lmax: int, x: np.ndarray) -> np.ndarray:
    """Fully normalised associated Legendre functions P[l, m, k] at sin(lat) x[k]."""
    c = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    P = np.zeros((lmax + 1, lmax + 1, x.size))
    P[0, 0] = 1.0
    for m in range(1, lmax + 1):
        P[m, m] = np.sqrt((2 * m + 1) / (2.0 * m)) * c * P[m - 1, m - 1]
    for m in range(0, lmax):
        P[m + 1, m] = np.sqrt(2 * m + 3.0) * x * P[m, m]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            a = np.sqrt((4.0 * l * l - 1) / (l * l - m * m))
            b = np.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1) ** 2 - 1))
            P[l, m] = a * (x * P[l - 1, m] - b * P[l - 2, m])
    return P


class SphericalField:
    """psi(lon, lat) = sum_lm P_lm(sin lat) (a_lm cos m lon + b_lm sin m lon)."""

    def __init__(self, lmax: int, seed: int, slope: float = -3.0):
        rng = np.random.default_rng(seed)
        self.lmax = lmax
        sd = np.zeros((lmax + 1, 1))
        sd[1:, 0] = np.arange(1, lmax + 1, dtype=float) ** (slope / 2)
        self.a = rng.standard_normal((lmax + 1, lmax + 1)) * sd
        self.b = rng.standard_normal((lmax + 1, lmax + 1)) * sd
        self.a[0] = 0.0
        self.b[0] = 0.0
        for l in range(lmax + 1):       # only m <= l, no sin(0 lon) term
            self.a[l, l + 1:] = 0.0
            self.b[l, l + 1:] = 0.0
            self.b[l, 0] = 0.0

    def _zonal(self, lat):
        P = legendre_table(self.lmax, np.sin(lat))          # (l, m, k)
        A = np.einsum("lm,lmk->mk", self.a, P)
        B = np.einsum("lm,lmk->mk", self.b, P)
        return A, B

    def eval(self, lon, lat, dlon_order=0):
        """Values (or d/dlon) on the grid lat (rows) x lon (columns)."""
        A, B = self._zonal(np.asarray(lat, float))
        m = np.arange(self.lmax + 1)[:, None]
        ml = m * np.asarray(lon, float)[None, :]
        if dlon_order == 0:
            return A.T @ np.cos(ml) + B.T @ np.sin(ml)
        return A.T @ (-m * np.sin(ml)) + B.T @ (m * np.cos(ml))

    def dlat(self, lon, lat, h=1e-6):
        lat = np.asarray(lat, float)
        return (self.eval(lon, lat + h) - self.eval(lon, lat - h)) / (2 * h)


def _lons(nlon):
    dlon = 2 * np.pi / nlon
    full = np.arange(nlon) * dlon
    return full, full + 0.5 * dlon, dlon


def winds_from_psi(psi_vtx, metrics, nlat):
    """Discretely non-divergent winds from vertex psi [nlat-1, nlon]."""
    le_lon = metrics[MESH["le_lon"]]
    le_lat = metrics[MESH["le_lat"], :nlat - 1]
    nlon = psi_vtx.shape[1]
    u = np.zeros((nlat, nlon))
    u[1:nlat - 1] = -(psi_vtx[1:] - psi_vtx[:-1]) / le_lon[1:nlat - 1, None]
    v = (psi_vtx - np.roll(psi_vtx, 1, axis=1)) / le_lat[:, None]
    return u, v


def balanced_geopotential(field: SphericalField, scale, metrics, nlat, nlon, omega, radius,
                          mean=5.0e4):
    """Phi with lap_h(Phi) = div_h(f grad psi) on the cell grid."""
    M = lambda k: metrics[MESH[k]]
    full, half, dlon = _lons(nlon)
    flat, hlat = M("full_lat"), M("half_lat")[:nlat - 1]
    # f grad(psi) on U points (x component) and V points (y component)
    fu = np.zeros((nlat, nlon))
    rows = slice(1, nlat - 1)
    dpsi_dlon = field.eval(half, flat[rows], 1) * scale
    fu[rows] = (2 * omega * np.sin(flat[rows]))[:, None] * dpsi_dlon / (
        radius * np.cos(flat[rows]))[:, None]
    fv = (2 * omega * np.sin(hlat))[:, None] * field.dlat(full, hlat) * scale / radius
    le_lon, le_lat = M("le_lon"), M("le_lat")[:nlat - 1]
    de_lon, de_lat = M("de_lon"), M("de_lat")[:nlat - 1]
    ac = M("area_cell")
    rhs = np.zeros((nlat, nlon))
    rhs[rows] = ((fu[rows] - np.roll(fu[rows], 1, axis=1)) * le_lon[rows, None]
                 + fv[1:] * le_lat[1:, None] - fv[:-1] * le_lat[:-1, None]) / ac[rows, None]
    rhs[0] = fv[0].sum() * le_lat[0] / nlon / ac[0]
    rhs[-1] = -fv[-1].sum() * le_lat[-1] / nlon / ac[-1]
    # Fourier in longitude; per wavenumber a tridiagonal system in latitude
    R = np.fft.rfft(rhs, axis=1) / nlon
    nm = R.shape[1]
    phat = np.zeros_like(R)
    cv = le_lat / de_lat                       # coupling across V row j
    for m in range(nm):
        lam = (2 * np.cos(m * dlon) - 2) * le_lon / np.where(de_lon > 0, de_lon, 1.0)
        if m == 0:
            n = nlat
            ab = np.zeros((3, n))
            diag = np.zeros(n)
            diag[1:-1] = -(cv[1:] + cv[:-1])
            diag[0], diag[-1] = -cv[0], -cv[-1]
            ab[1] = diag
            ab[0, 1:] = cv                    # super-diagonal
            ab[2, :-1] = cv                   # sub-diagonal
            b = (R[:, 0] * ac).astype(complex)
            # the constant is free: pin the south pole value
            ab[1, 0], ab[0, 1] = 1.0, 0.0
            b[0] = 0.0
            phat[:, 0] = solve_banded((1, 1), ab, b.real) + 0j
        else:
            n = nlat - 2
            ab = np.zeros((3, n))
            ab[1] = lam[1:-1] - (cv[1:] + cv[:-1])
            ab[0, 1:] = cv[1:-1]
            ab[2, :-1] = cv[1:-1]
            b = R[1:-1, m] * ac[1:-1]
            phat[1:-1, m] = (solve_banded((1, 1), ab, b.real)
                             + 1j * solve_banded((1, 1), ab, b.imag))
    phi = np.fft.irfft(phat * nlon, n=nlon, axis=1)
    w = ac[:, None] * np.ones((1, nlon))
    return phi - (phi * w).sum() / w.sum() + mean


def seeded_state(nlon, nlat, metrics, lmax, seed, vmax=40.0, omega=7.292123516990375e-05,
                 radius=6.37122e6, mean_gz=5.0e4, perturb=None):
    """(u_lon, v_lat, gz, gzs) of the seeded balanced recipe.  perturb =
    (rms_fraction, seed) adds an independent l**-3 field to psi scaled to that
    fraction of its rms (ensemble members)."""
    full, half, dlon = _lons(nlon)
    hlat = metrics[MESH["half_lat"], :nlat - 1]
    fld = SphericalField(lmax, seed)
    psi = fld.eval(half, hlat)
    if perturb is not None:
        frac, pseed = perturb
        p = SphericalField(lmax, pseed)
        dp = p.eval(half, hlat)
        psi_rms = np.sqrt((psi ** 2).mean())
        dp *= frac * psi_rms / np.sqrt((dp ** 2).mean())
        psi = psi + dp
        fld = _Sum(fld, p, frac * psi_rms / np.sqrt((p.eval(half, hlat) ** 2).mean()))
    u, v = winds_from_psi(psi, metrics, nlat)
    s = vmax / max(np.abs(u).max(), np.abs(v).max())
    u *= s
    v *= s
    gz = balanced_geopotential(fld, s, metrics, nlat, nlon, omega, radius, mean_gz)
    return u, v, gz, np.zeros((nlat, nlon))


class _Sum:
    """a + c * b for two SphericalFields (the perturbed member's psi)."""

    def __init__(self, a, b, c):
        self.a, self.b, self.c = a, b, c

    def eval(self, lon, lat, dlon_order=0):
        return self.a.eval(lon, lat, dlon_order) + self.c * self.b.eval(lon, lat, dlon_order)

    def dlat(self, lon, lat, h=1e-6):
        return self.a.dlat(lon, lat, h) + self.c * self.b.dlat(lon, lat, h)


def steady_geostrophic(nlon, nlat, metrics, omega, radius, alpha=0.0):
    """Williamson case 2 as steady_geostrophic_flow_test_set_ic (:20-67)."""
    full, half, _ = _lons(nlon)
    u0 = 2.0 * np.pi * radius / (12.0 * 86400.0)
    gz0 = 2.94e4
    ca, sa = np.cos(alpha), np.sin(alpha)
    fc, fs = metrics[MESH["full_cos_lat"]], metrics[MESH["full_sin_lat"]]
    u = u0 * (fc[:, None] * ca + np.cos(half)[None, :] * fs[:, None] * sa)
    v = np.broadcast_to(-u0 * np.cos(full)[None, :] * sa, (nlat - 1, nlon)).copy()
    gz = gz0 - (radius * omega * u0 + u0 ** 2 * 0.5) * (
        fs[:, None] * ca - np.cos(full)[None, :] * fc[:, None] * sa) ** 2
    return u, v, gz, np.zeros((nlat, nlon))


# ---- FFSL advection test inputs --------------------------------------------

ADV_PERIOD = 12 * 86400.0   # deform_test_mod.F90:22


def deform_wind_basis(nlon, nlat, metrics, lmax, seed, vmax=50.0):
    """(u_lon, v_lat) of the prescribed advection flow at full amplitude:
    discretely non-divergent winds from a seeded vertex streamfunction
    (l <= lmax, variance l**-3) scaled to a peak of vmax.  The flow at time t
    is this basis times cos(pi t / ADV_PERIOD), the time dependence of
    deform_case1_test_set_uv (deform_test_mod.F90:129-167).  Pole rows of u
    are zero: set_uv writes the rows between the poles only."""
    full, half, dlon = _lons(nlon)
    hlat = metrics[MESH["half_lat"], :nlat - 1]
    psi = SphericalField(lmax, seed).eval(half, hlat)
    u, v = winds_from_psi(psi, metrics, nlat)
    s = vmax / max(np.abs(u).max(), np.abs(v).max())
    return u * s, v * s


def wind_scales(nsteps, dt, time_step0=0, period=ADV_PERIOD):
    """Factor of the wind basis for adv_run at time steps time_step0 ..
    time_step0 + nsteps - 1: the driver sets the wind of step n at
    t = (n + 1) dt before calling adv_run (gmcore_adv_driver.F90:84-92)."""
    import math
    return np.array([math.cos(math.pi * ((n + 1) * dt) / period)
                     for n in range(time_step0, time_step0 + nsteps)])


def great_circle(lon1, lat1, lon, lat):
    """Central angle between (lon1, lat1) and the points (lon, lat)."""
    c = np.sin(lat1) * np.sin(lat) + np.cos(lat1) * np.cos(lat) * np.cos(lon - lon1)
    return np.arccos(np.clip(c, -1.0, 1.0))


def cosine_bells(nlon, nlat, metrics, seed, nbells=4, radius=0.5, background=0.1, amp=0.9):
    """Background plus nbells cosine bells amp * (1 + cos(pi r / radius)) / 2
    (r the central angle, radius 0.5 = R/2 on the sphere) at seeded centres
    uniform on the sphere; the bell shape of deform_test_set_ic
    (deform_test_mod.F90:75-90)."""
    rng = np.random.default_rng(seed)
    full, _, _ = _lons(nlon)
    lat = metrics[MESH["full_lat"]]
    LON, LAT = np.meshgrid(full, lat)
    q = np.full((nlat, nlon), background)
    for _ in range(nbells):
        lon0 = rng.uniform(0, 2 * np.pi)
        lat0 = np.arcsin(rng.uniform(-1, 1))
        r = great_circle(lon0, lat0, LON, LAT)
        q += np.where(r < radius, amp * 0.5 * (1 + np.cos(np.pi * r / radius)), 0.0)
    return q


def tracer_set(nlon, nlat, metrics, seed, ntracers):
    """ntracers - 1 seeded sums of 4 cosine bells and one uniform-random
    field in [0, 1], all positive."""
    qs = [cosine_bells(nlon, nlat, metrics, seed * 1000 + l) for l in range(ntracers - 1)]
    rng = np.random.default_rng(seed * 1000 + 999)
    qs.append(rng.uniform(0.0, 1.0, size=(nlat, nlon)))
    return qs
