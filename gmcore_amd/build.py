"""Build the native parts in-tree.

- libgmcore_hip.so: the gfx950 HIP library (hipcc), C ABI of include/gmcore_hip.h,
  include/gmcore_adv.h, include/gmcore_ens.h, include/gmcore_tlad.h,
  include/gmcore_io.h and include/gmcore_baro.h
- libgmcore_host.so: the Fortran host setup (amdflang): mesh metrics, filter tables
- gmcore_swm_gpu, gmcore_adv_gpu: the Fortran drivers (amdflang) on top of both

Every HIP translation unit is compiled for gfx950 with -ffp-contract=off
(each product and sum separately rounded, as the parity tests require) and
without packed-FP32 instruction selection.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["gmc_capi.hip", "gmc_ops.hip", "gmc_fused.hip", "gmc_adv.hip", "gmc_tlad.hip", "gmc_io.hip",
               "gmc_baro.hip"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]
F_FLAGS = ["-O2", "-fPIC", "-ffp-contract=off"]
F_HOST = ["gmc_kinds_mod.F90", "gmc_mesh_mod.F90", "gmc_filter_mod.F90", "gmc_lon_damp_mod.F90",
          "gmc_host_capi.F90"]
F_DRIVER = ["gmc_hip_mod.F90", "gmc_dynamics_mod.F90", "gmcore_swm_gpu.F90"]
F_ADV = ["gmc_adv_mod.F90", "gmcore_adv_gpu.F90"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "Default real literal" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return out


def build_hip() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libgmcore_hip.so")
    srcs = [os.path.join(HERE, "csrc", s) for s in HIP_SOURCES]
    tmp = out + ".tmp"
    _run([os.path.join(ROCM, "bin", "hipcc"), *HIP_FLAGS, "-shared", *srcs, "-o", tmp,
          f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, out)
    return out


def build_fortran() -> tuple[str, str]:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    fc = os.path.join(ROCM, "bin", "amdflang")
    objs = {}
    for src in F_HOST + F_DRIVER + F_ADV:
        obj = os.path.join(BUILD, src.replace(".F90", ".o"))
        _run([fc, *F_FLAGS, "-module-dir", BUILD, "-c", os.path.join(HERE, "fortran", src), "-o", obj])
        objs[src] = obj
    host = os.path.join(LIB, "libgmcore_host.so")
    _run([fc, *F_FLAGS, "-shared", *[objs[s] for s in F_HOST], "-o", host + ".tmp"])
    os.replace(host + ".tmp", host)
    drv = os.path.join(LIB, "gmcore_swm_gpu")
    _run([fc, *F_FLAGS, *[objs[s] for s in F_HOST[:3] + F_DRIVER], "-o", drv + ".tmp",
          "-L" + LIB, "-lgmcore_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(drv + ".tmp", drv)
    adv = os.path.join(LIB, "gmcore_adv_gpu")
    _run([fc, *F_FLAGS, *[objs[s] for s in F_HOST[:4] + F_DRIVER[:1] + F_ADV], "-o", adv + ".tmp",
          "-L" + LIB, "-lgmcore_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(adv + ".tmp", adv)
    return host, drv


def build_oracle() -> str:
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all() -> None:
    build_hip()
    build_fortran()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", os.listdir(LIB))
