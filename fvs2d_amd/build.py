"""Build libfvs2d_hip.so in place (hipcc for gfx950, g++ for host C++).

Run as ``python -m fvs2d_amd.build``.  The library lands next to this file so
it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfvs2d_hip.so")
BUILD = os.path.join(HERE, "_build")

ARCH = os.environ.get("FVS_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed-FP32 instructions are kept off (the solver is FP64; the network's
# float32 epilogues do not need them)
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
# The solver's kernels contract a multiply and an add into an FMA only within
# one expression (hipcc's default for HIP also fuses across statements, which
# the backend decides per use site): the same device function then rounds the
# same in every kernel it is inlined into, so the kernel arrangements give
# bit-identical states.
SOLVER_FLAGS = ["-ffp-contract=on"]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-variable"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    obj_host = os.path.join(BUILD, "mesh_prep.o")
    obj_dev = os.path.join(BUILD, "fvs_api.o")
    obj_mlp = os.path.join(BUILD, "mlp.o")
    cmds = [
        ["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "mesh_prep.cpp"), "-o", obj_host],
        [HIPCC, *HIP_FLAGS, *SOLVER_FLAGS, "-c", os.path.join(CSRC, "fvs_api.hip"), "-o", obj_dev],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "mlp.hip"), "-o", obj_mlp],
    ]
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        _run(c)
    tmp = LIB + ".tmp"
    # -z defs: a symbol left undefined is a link error here, not a load error on the GPU box
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-z,defs", obj_dev, obj_mlp, obj_host, "-o", tmp]
    if verbose:
        print(" ".join(link), flush=True)
    _run(link)
    os.replace(tmp, LIB)
    return LIB


FLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
FORTRAN = os.path.join(HERE, "fortran")
TESTS_F = os.path.join(os.path.dirname(HERE), "tests", "fortran")
FDRIVER = os.path.join(TESTS_F, "libfvs2d_fdriver.so")
# the reference is compiled with default REAL = 8 bytes (-r8)
F_FLAGS = ["-O2", "-fPIC", "-fdefault-real-8", "-fdefault-double-8"]


def build_fortran_harness(verbose: bool = False) -> str:
    """Compile the iso_c_binding modules (fvs2d_amd/fortran) with the test
    harness (tests/fortran) into tests/fortran/libfvs2d_fdriver.so."""
    if not os.path.exists(FLANG):
        return ""
    mod = os.path.join(BUILD, "fmod")
    os.makedirs(mod, exist_ok=True)
    srcs = [os.path.join(TESTS_F, "harness_modules.f90"), os.path.join(FORTRAN, "fvs2d_capi.f90"),
            os.path.join(FORTRAN, "fvs2d_dropin.f90"), os.path.join(TESTS_F, "driver.f90")]
    objs = []
    for src in srcs:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmd = [FLANG, *F_FLAGS, "-module-dir", mod, "-I", mod, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(obj)
    tmp = FDRIVER + ".tmp"
    link = [FLANG, "-shared", "-fPIC", *objs, "-L", HERE, "-lfvs2d_hip", f"-Wl,-rpath,{HERE}",
            "-Wl,-rpath,$ORIGIN/../../fvs2d_amd", "-o", tmp]
    if verbose:
        print(" ".join(link), flush=True)
    _run(link)
    os.replace(tmp, FDRIVER)
    return FDRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran_harness(verbose=True))
