"""Build the HIP library (and the Fortran driver) in-tree for gfx950.

    python qg-ybjp-model_amd/build.py [--jobs N] [--no-fortran]

Outputs:
    qg-ybjp-model_amd/lib/libqgybj.so     the C-ABI library (include/qgybj.h)
    qg-ybjp-model_amd/lib/qg_driver       the Fortran integration-loop driver
    qg-ybjp-model_amd/lib/libqgncio.so    its classic-netCDF reader/writer (host tests)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "lib", "obj")
INC = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
CC = shutil.which("gcc") or shutil.which("cc") or "/opt/rocm/llvm/bin/clang"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", INC, "-I", CSRC,
]

SOURCES = ["qg_step_kernels.hip", "qg_kernels.hip", "qg_adj_kernels.hip", "qg_adj_step.hip", "qg_api.cpp", "qg_ensemble.cpp", "qg_adj.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")
    # hipcc's host pass warns about the device-only target feature; keep quiet
    return r


def build_library(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    so = os.path.join(LIB, "libqgybj.so")
    tmp = so + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, so)
    return so


def build_ncio(verbose: bool = False) -> str:
    """The driver's classic-netCDF reader/writer (fortran/qg_ncio.c): an
    object for the driver and lib/libqgncio.so for the host tests."""
    os.makedirs(OBJ, exist_ok=True)
    src = os.path.join(FSRC, "qg_ncio.c")
    obj = os.path.join(OBJ, "qg_ncio.o")
    _run([CC, "-O2", "-std=c11", "-Wall", "-fPIC", "-c", src, "-o", obj], verbose)
    so = os.path.join(LIB, "libqgncio.so")
    _run([CC, "-shared", "-fPIC", obj, "-o", so + ".tmp"], verbose)
    os.replace(so + ".tmp", so)
    return obj


def build_driver(verbose: bool = False) -> str | None:
    """Compile the Fortran driver against libqgybj.so (amdflang)."""
    srcs = ["qg_params.f90", "qgybj_c.f90", "qg_ncio.f90", "qg_driver.f90"]
    if not all(os.path.exists(os.path.join(FSRC, s)) for s in srcs):
        return None
    ncio = build_ncio(verbose)
    exe = os.path.join(LIB, "qg_driver")
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    cmd = [FC, "-O2", "-J", mod, *[os.path.join(FSRC, s) for s in srcs], ncio,
           "-L", LIB, "-lqgybj", f"-Wl,-rpath,{LIB}", "-Wl,-rpath,$ORIGIN", "-o", exe]
    _run(cmd, verbose)
    return exe


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--no-fortran", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    so = build_library(a.jobs, a.verbose)
    print("built", so)
    if not a.no_fortran:
        exe = build_driver(a.verbose)
        if exe:
            print("built", exe)


if __name__ == "__main__":
    main()
