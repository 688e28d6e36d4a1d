"""The C-ABI libraries load and export every entry point their headers
declare (CPU only: no device calls)."""
import ctypes as C
import os
import re

import pytest

from gmcore_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gmc_[a-z_0-9]+)\s*\(", text)))


HEADERS = ["gmcore_hip.h", "gmcore_adv.h", "gmcore_ens.h", "gmcore_tlad.h", "gmcore_io.h", "gmcore_baro.h"]


def test_every_header_is_checked():
    assert sorted(HEADERS) == sorted(f for f in os.listdir(os.path.join(ROOT, "include"))
                                     if f.endswith(".h"))


@pytest.mark.parametrize("header", HEADERS)
def test_hip_library_exports_header(header):
    names = declared(header)
    assert len(names) >= 6
    lib = C.CDLL(capi.LIB_HIP)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("header", HEADERS)
def test_python_mirror_covers_header(header):
    lib = capi.load_hip()
    for n in declared(header):
        assert getattr(lib, n).argtypes is not None, f"{n} has no ctypes signature in capi.py"


def test_host_library_exports():
    lib = C.CDLL(capi.LIB_HOST)
    for n in ("gmch_mesh_metrics", "gmch_filter", "gmch_lon_damp_weights"):
        assert hasattr(lib, n)


def test_struct_layouts_match_header():
    # gmc_mesh: int32, int32, double, pointer; gmc_filter: int32 + 4 pointers
    assert C.sizeof(capi.GmcMesh) == 24
    assert C.sizeof(capi.GmcFilter) == 40
    assert capi.GmcConfig.big_filter.offset % 8 == 0
    assert capi.GMC_MESH_NFIELDS == 32
    # gmc_adv_config: 2 doubles, 6 int32, 1 pointer
    assert C.sizeof(capi.GmcAdvConfig) == 48
    assert capi.GmcAdvConfig.lon_damp_wgt.offset == 40


def test_fortran_bindings_name_exported_symbols():
    """Every bind(C, name=...) interface of the Fortran host module names a
    function the library exports."""
    text = open(os.path.join(ROOT, "gmcore_amd", "fortran", "gmc_hip_mod.F90")).read()
    names = re.findall(r"bind\(C, name='(gmc_[a-z_0-9]+)'\)", text)
    assert len(names) > 40
    lib = C.CDLL(capi.LIB_HIP)
    assert not [n for n in names if not hasattr(lib, n)]


def test_fortran_baro_field_codes_match_header():
    """The GMB_* parameters of the Fortran module are the enum of
    include/gmcore_baro.h, in order."""
    text = open(os.path.join(ROOT, "gmcore_amd", "fortran", "gmc_hip_mod.F90")).read()
    fortran = {k: int(v) for k, v in re.findall(r"(GMB_[A-Z_]+) = (\d+)", text)}
    header = open(os.path.join(ROOT, "include", "gmcore_baro.h")).read()
    enum = re.search(r"enum gmc_baro_field \{(.*?)\};", header, re.S).group(1)
    order = [m for m in re.findall(r"\b(GMB_[A-Z_]+)\b", re.sub(r"/\*.*?\*/", "", enum)) if m != "GMB_NFIELDS"]
    assert fortran == {name: k for k, name in enumerate(order)}
    assert list(capi.BARO_FIELD) == [n[4:].lower() for n in order]
