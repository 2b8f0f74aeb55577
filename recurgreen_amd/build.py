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
SOURCES = ["rg_plan.hip", "rg_svd.hip", "rg_kernels_f64.hip", "rg_blocked.hip", "rg_gj.hip", "rg_project.hip", "rg_eig.hip", "rg_zgemm.hip", "rg_refl.hip", "rg_adjoint.hip", "rg_nopiv.hip", "rg_dyson32.hip", "rg_schur.hip", "rg_tprofile.hip", "rg_dyson64.hip", "rg_dyson_bf3.hip", "rg_noalloc.hip", "rg_nopiv64.hip", "rg_project_gen.hip", "rg_refine.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops is switched off for every HIP object: the v_pk_*_f32 forms are avoided.
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I", os.path.join(HERE, "..", "include"),
]


# Per-source additions: the generating projection keeps its generator's FP64 constants inside the chunk loop
# (MachineLICM would hoist them into 88 registers that its accumulators need).
EXTRA_FLAGS = {"rg_project_gen.hip": ["-mllvm", "-disable-machine-licm"]}


def build(verbose=False, only_changed=False):
    """Compile every HIP source for gfx950 (one hipcc per file, in parallel) and link the
    shared library in-tree; raises on failure.  only_changed: recompile only the sources newer than
    their object (or than rg_internal.h / the ABI header)."""
    from concurrent.futures import ThreadPoolExecutor

    odir = os.path.join(HERE, "_build")
    os.makedirs(odir, exist_ok=True)
    objs = [os.path.join(odir, os.path.splitext(f)[0] + ".o") for f in SOURCES]

    hdrs = [os.path.join(CSRC, "rg_internal.h"), os.path.join(CSRC, "rg_cos.h"),
            os.path.join(HERE, "..", "include", "recurgreen_hip.h")]
    newest_hdr = max(os.path.getmtime(h) for h in hdrs)

    def stale(i):
        src = os.path.join(CSRC, SOURCES[i])
        return (not only_changed or not os.path.exists(objs[i])
                or os.path.getmtime(objs[i]) < max(os.path.getmtime(src), newest_hdr))

    def compile_one(i):
        if not stale(i):
            return None, None
        cmd = [HIPCC, f"--offload-arch={ARCH}", *FLAGS, *EXTRA_FLAGS.get(SOURCES[i], []), "-c",
               os.path.join(CSRC, SOURCES[i]), "-o", objs[i]]
        if verbose:
            print(" ".join(cmd))
        return cmd, subprocess.run(cmd, capture_output=True, text=True)

    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        results = list(ex.map(compile_one, range(len(SOURCES))))
    for cmd, res in results:
        if res is not None and res.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed compiling {os.path.basename(cmd[-3])}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking librecurgreen_hip.so")
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
    quick = "--changed" in sys.argv
    print(build(verbose=True, only_changed=quick))
    print(build_fortran(verbose=True))
