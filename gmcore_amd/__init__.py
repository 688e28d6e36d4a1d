"""MI355X (gfx950) implementation of GMCORE's shallow-water dynamics path.

The compute lives in libgmcore_hip.so (HIP kernels behind the C ABI of
include/gmcore_hip.h); mesh and filter set-up in the Fortran host library
libgmcore_host.so.  This package is the Python host over both.
"""
from .namelist import Namelist, parse_namelist  # noqa: F401
