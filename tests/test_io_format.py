"""h0 and r0 files (include/gmcore_io.h) in the netCDF classic format, read
back with scipy's independent netCDF-3 reader: the reference's dimensions
(time unlimited, lon, lat, ilon, ilat), variable names, dimension order,
units and long names (history_mod.F90:125-150, restart_mod.F90:54-149), the
values bit for bit (4-byte floats for tm and zs, whose dtype the reference
leaves to its I/O library's default), records appended in order, and the
r0 round trip through gmc_r0_read.  CPU only: the host functions of the
library need no device."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from gmcore_amd import io
from gmcore_amd.host import Mesh
from gmcore_amd.namelist import Namelist

NLON, NLAT = 36, 19


def _coords():
    nl = Namelist()
    m = Mesh(NLON, NLAT, nl.radius, nl.omega)
    lon, ilon = io.lon_degrees(NLON, m.dlon)
    return lon, m["full_lat_deg"], ilon, m["half_lat_deg"][:NLAT - 1]


def _fields(seed):
    r = np.random.default_rng(seed)
    return {"u": r.standard_normal((NLAT, NLON)), "v": r.standard_normal((NLAT - 1, NLON)),
            "z": r.standard_normal((NLAT, NLON)) * 1e3, "pv": r.standard_normal((NLAT - 1, NLON)) * 1e-8,
            "div": r.standard_normal((NLAT, NLON)) * 1e-6, "zs": r.standard_normal((NLAT, NLON))}


def test_h0_layout_and_records(tmp_path):
    path = str(tmp_path / "case.h0.nc")
    c = _coords()
    recs = [_fields(1), _fields(2), _fields(3)]
    tot = [(1.5e18, 6.5e22, 3.7e2), (1.5e18 + 7, 6.4e22, 3.6e2), (1.5e18 + 9, 6.3e22, 3.5e2)]
    for k, f in enumerate(recs):
        io.write_h0(path, f, c, 0.25 * k, "days since 2000-01-01 00:00:00", 600.0, tot[k], new_file=k == 0)
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.version_byte == 2
        assert nc.dimensions["time"] is None
        assert (nc.dimensions["lon"], nc.dimensions["lat"], nc.dimensions["ilon"], nc.dimensions["ilat"]) == \
            (NLON, NLAT, NLON, NLAT - 1)
        assert nc.time_step_size == 600.0
        v = nc.variables
        assert v["u"].dimensions == ("time", "lat", "ilon")
        assert v["v"].dimensions == ("time", "ilat", "lon")
        assert v["z"].dimensions == ("time", "lat", "lon")
        assert v["pv"].dimensions == ("time", "ilat", "ilon")
        assert v["div"].dimensions == ("time", "lat", "lon")
        assert v["zs"].dimensions == ("lat", "lon")
        assert v["tm"].dimensions == ("time",)
        assert v["tm"].typecode() == "f" and v["zs"].typecode() == "f" and v["te"].typecode() == "d"
        assert v["u"].units == b"m s-1" and v["pv"].long_name == b"potential vorticity"
        assert v["time"].units == b"days since 2000-01-01 00:00:00"
        assert np.array_equal(v["time"][:], [0.0, 0.25, 0.5])
        for k, f in enumerate(recs):
            for name in ("u", "v", "z", "pv", "div"):
                assert np.array_equal(v[name][k], f[name]), (k, name)
            assert v["te"][k] == tot[k][1] and v["tpe"][k] == tot[k][2]
            assert v["tm"][k] == np.float32(tot[k][0])
        assert np.array_equal(v["zs"][:], recs[0]["zs"].astype(np.float32))
        assert np.array_equal(v["lat"][:], c[1]) and np.array_equal(v["ilon"][:], c[2])


def test_r0_round_trip(tmp_path):
    path = str(tmp_path / "case.r0.nc")
    f = _fields(4)
    gz = f["z"] * 9.8
    io.write_r0(path, f["u"], f["v"], gz, f["zs"], _coords(), 86400.0, "seconds since 2000-01-01 00:00:00",
                600.0, "1 days")
    u, v, g2, gzs, t = io.read_r0(path, NLON, NLAT)
    assert np.array_equal(u, f["u"]) and np.array_equal(v, f["v"]) and np.array_equal(g2, gz)
    assert np.array_equal(gzs, f["zs"]) and t == 86400.0
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.restart_interval == b"1 days"
        assert nc.variables["gz"].dimensions == ("time", "lat", "lon")
        assert nc.variables["gz"].units == b"m2 s-2"
        assert np.array_equal(nc.variables["gzs"][0], f["zs"])


def test_r0_read_refuses_wrong_grid_and_other_files(tmp_path):
    path = str(tmp_path / "case.r0.nc")
    f = _fields(5)
    io.write_r0(path, f["u"], f["v"], f["z"], f["zs"], _coords(), 0.0, "s", 600.0, "1 days")
    with pytest.raises(OSError):
        io.read_r0(path, NLON * 2, NLAT)
    junk = tmp_path / "junk.nc"
    junk.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))     # a netCDF-4 (HDF5) signature
    with pytest.raises(OSError):
        io.read_r0(str(junk), NLON, NLAT)
