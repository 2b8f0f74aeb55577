"""Build the gfx950 HIP library librecurgreen_hip.so in-tree with hipcc.

Usage: python -m recurgreen_amd.build   (or recurgreen_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librecurgreen_hip.so")
SOURCES = ["rg_plan.hip", "rg_kernels_f64.hip", "rg_blocked.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops is switched off for every HIP object: the v_pk_*_f32 forms are avoided.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building librecurgreen_hip.so")
    os.replace(tmp, LIB)
    return LIB


FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FEXE = os.path.join(FDIR, "recurgreen_gpu")


def build_fortran(verbose: bool = False) -> str:
    """Fortran host (iso_c_binding module + namelist driver) linked against the HIP library."""
    bdir = os.path.join(FDIR, "_build")
    os.makedirs(bdir, exist_ok=True)
    steps = [
        [FC, "-O2", "-J", bdir, "-c", os.path.join(FDIR, "recurgreen_hip.f90"), "-o", os.path.join(bdir, "rh.o")],
        [FC, "-O2", "-J", bdir, "-c", os.path.join(FDIR, "recurgreen_gpu.f90"), "-o", os.path.join(bdir, "rg.o")],
        [FC, "-o", FEXE + ".tmp", os.path.join(bdir, "rh.o"), os.path.join(bdir, "rg.o"), "-L", HERE,
         "-lrecurgreen_hip", "-Wl,-rpath,$ORIGIN/.."],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("amdflang failed building the Fortran host")
    os.replace(FEXE + ".tmp", FEXE)
    return FEXE


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
