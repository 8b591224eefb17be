"""Builds the native parts of fdtomoc_amd in-tree.

    python -m fdtomoc_amd.build

produces ``fdtomoc_amd/libfdtomoc_amd.so`` (gfx950 kernels -- the sphfd sweep,
c2f's coarse-to-fine interpolation, sphfdloc's location and sphrayderv's ray
tracing -- + C ABI of include/fdtomoc_amd.h) and ``fdtomoc_amd/bin/sphfd`` (the drop-in
executable).  All code is compiled without floating-point contraction: the
FP64 stencils must round like the reference's x86-64 build.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfdtomoc_amd.so")
EXE = os.path.join(PKG, "bin", "sphfd")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("FDTOMOC_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Wno-unknown-pragmas"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [("sweep_kernel.hip", True), ("coarse_kernel.hip", True), ("fdgpu.cpp", False), ("sphfd_io.cpp", False),
           ("table_prep.cpp", False), ("sphfd_drop.cpp", False), ("serve.cpp", False), ("coarse.cpp", False),
           ("loc_kernel.hip", True), ("loc.cpp", False), ("table_io.cpp", False),
           ("table_writer.cpp", False), ("loc_geo.cpp", False), ("ray_kernel.hip", True), ("ray.cpp", False)]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(EXE), exist_ok=True)

    def compile_one(item):
        src, dev = item
        obj = os.path.join(BUILD, src + ".o")
        host = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", f"-I{os.path.join(ROCM, 'include')}"]
        cmd = [HIPCC] + COMMON + (DEVICE if dev else host) + ["-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd, verbose)
        return obj

    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    exe_tmp = EXE + ".tmp"
    _run(["gcc", "-O2", "-o", exe_tmp, os.path.join(CSRC, "sphfd_main.c"), f"-L{PKG}", "-lfdtomoc_amd",
          "-Wl,-rpath,$ORIGIN/.."], verbose)
    os.replace(exe_tmp, EXE)
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
