"""ctypes mirror of include/gmcore_hip.h, include/gmcore_adv.h,
include/gmcore_ens.h and include/gmcore_tlad.h (structs, enums, signatures).

The C ABI is the drop-in boundary of the dynamics path; this module only
describes it.  ``load_hip()`` opens the in-tree ``libgmcore_hip.so`` and fails
loudly when it has not been built -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_HIP = os.path.join(HERE, "lib", "libgmcore_hip.so")
LIB_HOST = os.path.join(HERE, "lib", "libgmcore_host.so")

# error codes
GMC_OK, GMC_EINVAL, GMC_EDEVICE, GMC_ENOMEM, GMC_ENUMERIC, GMC_ESTATE = range(6)

MESH_FIELDS = [
    "full_lat", "half_lat", "full_lat_deg", "half_lat_deg",
    "full_cos_lat", "half_cos_lat", "full_sin_lat", "half_sin_lat",
    "area_cell", "area_lon", "area_lon_west", "area_lon_east",
    "area_lon_north", "area_lon_south",
    "area_lat", "area_lat_west", "area_lat_east", "area_lat_north",
    "area_lat_south", "area_vtx", "area_subcell1", "area_subcell2",
    "de_lon", "le_lon", "de_lat", "le_lat",
    "full_f", "half_f",
    "full_tangent_wgt1", "full_tangent_wgt2",
    "half_tangent_wgt1", "half_tangent_wgt2",
]
MESH = {name: k for k, name in enumerate(MESH_FIELDS)}
GMC_MESH_NFIELDS = len(MESH_FIELDS)

GMC_TIME_PC2, GMC_TIME_WRFRK3, GMC_TIME_EULER = 0, 1, 2
GMC_PV_MIDPOINT, GMC_PV_UPWIND, GMC_PV_TVD = 0, 1, 2
GMC_TVD_NONE, GMC_TVD_VAN_LEER, GMC_TVD_MC = 0, 1, 2
PV_SCHEMES = {"midpoint": GMC_PV_MIDPOINT, "upwind": GMC_PV_UPWIND, "tvd": GMC_PV_TVD}
TVD_LIMITERS = {"none": GMC_TVD_NONE, "van_leer": GMC_TVD_VAN_LEER, "mc": GMC_TVD_MC}

FIELDS = ["u_lon", "v_lat", "gz", "u_f", "v_f", "gz_f", "gzs", "du", "dv",
          "dgz", "pv", "ke", "div", "m"]
FIELD = {name: k for k, name in enumerate(FIELDS)}
# grid of each field: "U" (nlat rows), "V" (nlat-1), "C" (nlat), "X" (nlat-1)
FIELD_GRID = {"u_lon": "U", "v_lat": "V", "gz": "C", "u_f": "U", "v_f": "V",
              "gz_f": "C", "gzs": "C", "du": "U", "dv": "V", "dgz": "C",
              "pv": "X", "ke": "C", "div": "C", "m": "C"}


def field_rows(name: str, nlat: int) -> int:
    return nlat - 1 if FIELD_GRID[name] in ("V", "X") else nlat


class GmcMesh(C.Structure):
    _fields_ = [("nlon", C.c_int32), ("nlat", C.c_int32), ("dlon", C.c_double),
                ("metrics", C.POINTER(C.c_double))]


class GmcFilter(C.Structure):
    _fields_ = [("max_taps", C.c_int32),
                ("ngrid_lon", C.POINTER(C.c_int32)), ("wgt_lon", C.POINTER(C.c_double)),
                ("ngrid_lat", C.POINTER(C.c_int32)), ("wgt_lat", C.POINTER(C.c_double))]


class GmcConfig(C.Structure):
    _fields_ = [("g", C.c_double), ("time_scheme", C.c_int32), ("pv_scheme", C.c_int32),
                ("upwind_order_pv", C.c_int32), ("upwind_wgt_pv", C.c_double),
                ("pv_pole_stokes", C.c_int32), ("coriolis_scheme", C.c_int32),
                ("use_div_damp", C.c_int32), ("div_damp_order", C.c_int32),
                ("div_damp_coef2", C.c_double),
                ("big_filter", GmcFilter), ("small_filter1", GmcFilter),
                ("div_damp_coef4", C.c_double), ("dt_dyn", C.c_double),
                ("tvd_limiter", C.c_int32), ("reserved", C.c_int32)]


GMC_LIMITER_MONO, GMC_LIMITER_PD, GMC_LIMITER_NONE = 0, 1, 2
LIMITERS = {"mono": GMC_LIMITER_MONO, "pd": GMC_LIMITER_PD, "none": GMC_LIMITER_NONE}
GMC_FLUX_PPM, GMC_FLUX_VAN_LEER = 0, 1
FLUX_TYPES = {"ppm": GMC_FLUX_PPM, "van_leer": GMC_FLUX_VAN_LEER}


class GmcH0Record(C.Structure):
    """gmc_h0_record of include/gmcore_io.h"""
    _fields_ = [("nlon", C.c_int32), ("nlat", C.c_int32),
                ("lon_deg", C.POINTER(C.c_double)), ("lat_deg", C.POINTER(C.c_double)),
                ("ilon_deg", C.POINTER(C.c_double)), ("ilat_deg", C.POINTER(C.c_double)),
                ("time_value", C.c_double), ("time_units", C.c_char_p), ("dt", C.c_double),
                ("tm", C.c_double), ("te", C.c_double), ("tpe", C.c_double),
                ("u", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)), ("z", C.POINTER(C.c_double)),
                ("pv", C.POINTER(C.c_double)), ("div", C.POINTER(C.c_double)), ("zs", C.POINTER(C.c_double))]


class GmcVcoord(C.Structure):
    """gmc_vcoord of include/gmcore_baro.h"""
    _fields_ = [("nlev", C.c_int32), ("reserved", C.c_int32), ("p0", C.c_double), ("ptop", C.c_double),
                ("hyai", C.POINTER(C.c_double)), ("hybi", C.POINTER(C.c_double))]


class GmcThermo(C.Structure):
    """gmc_thermo of include/gmcore_baro.h"""
    _fields_ = [("Rd", C.c_double), ("cpd", C.c_double), ("p0", C.c_double)]


# gmc_baro_field: (levels: 0 one, 1 full, 2 half; V-point rows)
BARO_FIELDS = {
    "phs": (0, False), "gzs": (0, False), "dphs": (0, False), "pt": (1, False), "ph_lev": (2, False),
    "ph_exn_lev": (2, False), "ph": (1, False), "t": (1, False), "gz_lev": (2, False), "gz": (1, False),
    "dmfdlon": (1, False), "dmfdlat": (1, False), "we_lev": (2, False), "we_lev_lon": (2, False),
    "we_lev_lat": (2, True), "u_f": (1, False), "v_f": (1, True), "m_lon": (1, False), "m_lat": (1, True),
    "wedudlev": (1, False), "wedvdlev": (1, True), "pgf_lon": (1, False), "pgf_lat": (1, True),
    "cflz": (2, False), "ptf_lev": (2, False), "dptfdlev": (1, False), "u_lon": (1, False), "v_lat": (1, True),
    "adv_u": (1, False), "adv_v": (1, True), "adv_mfx": (1, False), "adv_mfy": (1, True),
    "adv_cflx": (1, False), "adv_cfly": (1, True), "adv_divx": (1, False), "adv_divy": (1, False),
    "ptf_lon": (1, False), "ptf_lat": (1, True), "dptfdlon": (1, False), "dptfdlat": (1, False),
}
BARO_FIELD = {name: k for k, name in enumerate(BARO_FIELDS)}


class GmcAdvConfig(C.Structure):
    """gmc_adv_config of include/gmcore_adv.h"""
    _fields_ = [("radius", C.c_double), ("dt_adv", C.c_double), ("nstep", C.c_int32),
                ("ntracers", C.c_int32), ("lon_halo_width", C.c_int32), ("limiter", C.c_int32),
                ("lon_damp_passes", C.c_int32), ("flux_type", C.c_int32),
                ("lon_damp_wgt", C.POINTER(C.c_double))]


def dptr(a):
    """double* of a contiguous float64 numpy array."""
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


_hip = None


def load_hip() -> C.CDLL:
    """Open libgmcore_hip.so (built by __graft_entry__.build()).  Raises if absent."""
    global _hip
    if _hip is not None:
        return _hip
    if not os.path.exists(LIB_HIP):
        raise RuntimeError(f"{LIB_HIP} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(LIB_HIP)
    vp = C.c_void_p
    P = C.POINTER
    sig = {
        "gmc_create": (C.c_int, [P(GmcMesh), P(GmcConfig), P(vp)]),
        "gmc_destroy": (C.c_int, [vp]),
        "gmc_last_error": (C.c_char_p, [vp]),
        "gmc_upload": (C.c_int, [vp, C.c_int32, C.c_int32, P(C.c_double), C.c_int64]),
        "gmc_download": (C.c_int, [vp, C.c_int32, C.c_int32, P(C.c_double), C.c_int64]),
        "gmc_prepare": (C.c_int, [vp, C.c_int32]),
        "gmc_space_operators": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_double, C.c_int32]),
        "gmc_update_state": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32]),
        "gmc_step": (C.c_int, [vp, C.c_double, C.c_int32]),
        "gmc_time_integrate": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_double]),
        "gmc_current_level": (C.c_int, [vp, P(C.c_int32)]),
        "gmc_diagnose": (C.c_int, [vp, C.c_int32, P(C.c_double)]),
        "gmc_synchronize": (C.c_int, [vp]),
        "gmc_kernel_times": (C.c_int, [vp, C.c_double, C.c_int32, P(C.c_double), C.c_int32,
                                       P(C.c_int32)]),
        "gmc_kernel_name": (C.c_char_p, [C.c_int32]),
        "gmc_stream": (C.c_int, [vp, P(vp)]),
        "gmc_field_ptr": (C.c_int, [vp, C.c_int32, C.c_int32, P(P(C.c_double)), P(C.c_int64)]),
        "gmc_filter_plan": (C.c_int, [vp, P(C.c_int32)]),
        "gmc_keep_derived": (C.c_int, [vp, C.c_int32]),
        "gmc_diagnose_dev": (C.c_int, [vp, C.c_int32, vp]),
        # FFSL advection, include/gmcore_adv.h
        "gmc_adv_create": (C.c_int, [P(GmcMesh), P(GmcAdvConfig), P(vp)]),
        "gmc_adv_destroy": (C.c_int, [vp]),
        "gmc_adv_last_error": (C.c_char_p, [vp]),
        "gmc_adv_upload_tracer": (C.c_int, [vp, C.c_int32, C.c_int32, P(C.c_double), C.c_int64]),
        "gmc_adv_download_tracer": (C.c_int, [vp, C.c_int32, C.c_int32, P(C.c_double), C.c_int64]),
        "gmc_adv_set_wind": (C.c_int, [vp, P(C.c_double), P(C.c_double), P(C.c_double),
                                       P(C.c_double), P(C.c_double), C.c_int64]),
        "gmc_adv_run": (C.c_int, [vp, C.c_int64]),
        "gmc_adv_set_wind_basis": (C.c_int, [vp, P(C.c_double), P(C.c_double), C.c_int64]),
        "gmc_adv_run_dyn": (C.c_int, [vp, vp, C.c_int32, C.c_int64]),
        "gmc_adv_init_dyn": (C.c_int, [vp, vp, C.c_int32]),
        "gmc_adv_drive": (C.c_int, [vp, C.c_int64, P(C.c_double), C.c_int32]),
        "gmc_adv_old_level": (C.c_int, [vp, P(C.c_int32)]),
        "gmc_adv_mass": (C.c_int, [vp, C.c_int32, P(C.c_double)]),
        "gmc_adv_fix_count": (C.c_int, [vp, P(C.c_int64)]),
        "gmc_adv_synchronize": (C.c_int, [vp]),
        "gmc_adv_stream": (C.c_int, [vp, P(vp)]),
        "gmc_adv_kernel_times": (C.c_int, [vp, C.c_int64, P(C.c_double), C.c_int32, P(C.c_double),
                                           C.c_int32, P(C.c_int32)]),
        "gmc_adv_kernel_name": (C.c_char_p, [C.c_int32]),
        "gmc_adv_tl_drive": (C.c_int, [vp, C.c_int64, P(C.c_double), C.c_int32, P(C.c_double), C.c_int64]),
        "gmc_adv_ad_drive": (C.c_int, [vp, C.c_int64, P(C.c_double), C.c_int32, C.c_int32, P(C.c_double),
                                       C.c_int64]),
        "gmc_adv_gradient": (C.c_int, [vp, C.c_int64, P(C.c_double), C.c_int32, C.c_int32, P(C.c_double),
                                       P(C.c_double), P(C.c_double), C.c_int64]),
        # history and restart files, include/gmcore_io.h
        "gmc_h0_write": (C.c_int, [C.c_char_p, P(GmcH0Record), C.c_int32]),
        "gmc_r0_write": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32] + [P(C.c_double)] * 4
                         + [C.c_double, C.c_char_p, C.c_double, C.c_char_p] + [P(C.c_double)] * 4),
        "gmc_r0_read": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32] + [P(C.c_double)] * 5),
        "gmc_history_h0": (C.c_int, [vp, C.c_char_p, C.c_int32, C.c_double, C.c_char_p, C.c_double, C.c_int32]),
        "gmc_restart_write": (C.c_int, [vp, C.c_char_p, C.c_int32, C.c_double, C.c_char_p, C.c_double,
                                        C.c_char_p]),
        "gmc_restart_read": (C.c_int, [vp, C.c_char_p, C.c_int32, P(C.c_double)]),
        # hydrostatic column operators, include/gmcore_baro.h
        "gmc_baro_create": (C.c_int, [P(GmcMesh), P(GmcVcoord), P(GmcThermo), P(vp)]),
        "gmc_baro_destroy": (C.c_int, [vp]),
        "gmc_baro_last_error": (C.c_char_p, [vp]),
        "gmc_baro_set": (C.c_int, [vp, C.c_int32, P(C.c_double), C.c_int64]),
        "gmc_baro_get": (C.c_int, [vp, C.c_int32, P(C.c_double), C.c_int64]),
        "gmc_baro_calc_ph": (C.c_int, [vp]),
        "gmc_baro_calc_t": (C.c_int, [vp]),
        "gmc_baro_calc_gz_lev": (C.c_int, [vp]),
        "gmc_baro_calc_dphsdt": (C.c_int, [vp]),
        "gmc_baro_calc_we_lev": (C.c_int, [vp]),
        "gmc_baro_calc_wedudlev_wedvdlev": (C.c_int, [vp]),
        "gmc_baro_pgf_lin97": (C.c_int, [vp]),
        "gmc_baro_calc_ptf_lev": (C.c_int, [vp, C.c_int32]),
        "gmc_baro_calc_ptf_hflx": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_double]),
        "gmc_baro_calc_grad_ptf": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_double]),
        "gmc_baro_set_small_filter2": (C.c_int, [vp, P(GmcFilter)]),
        "gmc_baro_small_filter2": (C.c_int, [vp]),
        "gmc_baro_synchronize": (C.c_int, [vp]),
        "gmc_baro_stream": (C.c_int, [vp, P(vp)]),
        # batched ensembles, include/gmcore_ens.h
        "gmc_ens_create": (C.c_int, [P(vp), C.c_int32, P(vp)]),
        "gmc_ens_destroy": (C.c_int, [vp]),
        "gmc_ens_last_error": (C.c_char_p, [vp]),
        "gmc_ens_step": (C.c_int, [vp, P(C.c_double), C.c_int32]),
        "gmc_ens_autotune": (C.c_int, [vp, P(C.c_double), C.c_int32, P(C.c_int32), P(C.c_double),
                                       C.c_int32, P(C.c_int32)]),
        "gmc_ens_shape": (C.c_int, [vp, C.c_int32]),
        "gmc_ens_shape_name": (C.c_char_p, [C.c_int32]),
        "gmc_ens_stream": (C.c_int, [vp, P(vp)]),
        # tangent-linear and adjoint, include/gmcore_tlad.h
        "gmc_tlad_create": (C.c_int, [vp, P(vp)]),
        "gmc_tlad_destroy": (C.c_int, [vp]),
        "gmc_tlad_last_error": (C.c_char_p, [vp]),
        "gmc_tlad_upload": (C.c_int, [vp, C.c_int32, C.c_int32, P(C.c_double), C.c_int64]),
        "gmc_tlad_download": (C.c_int, [vp, C.c_int32, C.c_int32, P(C.c_double), C.c_int64]),
        "gmc_tlad_zero": (C.c_int, [vp]),
        "gmc_tl_step": (C.c_int, [vp, C.c_double, C.c_int32]),
        "gmc_ad_step": (C.c_int, [vp, C.c_double, C.c_int32]),
        "gmc_gradient": (C.c_int, [vp, C.c_double, C.c_int32, C.c_int32, C.c_int32, P(C.c_int64),
                                   P(C.c_double), C.c_double, P(C.c_double), P(C.c_double),
                                   P(C.c_double), C.c_double, C.c_int64, P(C.c_double), P(C.c_double),
                                   P(C.c_double), P(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    _hip = lib
    return lib
