"""Builds libbddcml_amd.so (HIP for gfx950 + host C++) in-tree.

    python -m bddcml_amd.build

Every HIP source is compiled with hipcc for gfx950 only; packed FP32 ops are
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).  Matrix-core
instructions use the VGPR form for their accumulators
(-mllvm -amdgpu-mfma-vgpr-form): with accumulator registers the compiler
copied the loop-carried accumulators between register files on every
unrolled step of the tile products and drained the matrix pipeline there.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbddcml_amd.so")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

SOURCES = ["kernels_tile.hip", "kernels_misc.hip", "device.hip", "adapt.hip", "noalloc.hip", "constraints.cpp", "plan.cpp",
           "levels.cpp", "global.cpp", "api.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function"]
HIPFLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-mllvm", "-amdgpu-mfma-vgpr-form"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_LIB = os.path.join(FORTRAN_DIR, "libbddcml_amd_fortran.so")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


def build_fortran(verbose: bool = False) -> str:
    """The Fortran interface (module module_bddcml over the C entry points):
    its module file goes to fortran/include, its code into
    libbddcml_amd_fortran.so, linked against libbddcml_amd.so.  A Fortran
    caller compiles with -I bddcml_amd/fortran/include and links both."""
    inc = os.path.join(FORTRAN_DIR, "include")
    os.makedirs(inc, exist_ok=True)
    obj = os.path.join(inc, "module_bddcml.o")
    out = _run([FC, "-O2", "-fPIC", "-module-dir", inc, "-c",
                os.path.join(FORTRAN_DIR, "module_bddcml.f90"), "-o", obj])
    if verbose and out:
        print(out)
    tmp = FORTRAN_LIB + ".tmp"
    _run([FC, "-shared", "-o", tmp, obj, "-L" + HERE, "-lbddcml_amd", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, FORTRAN_LIB)
    os.remove(obj)
    return FORTRAN_LIB


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_objs")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC] + COMMON + HIPFLAGS + ["-c", path, "-o", obj]
        if src.endswith(".cpp"):
            # host plan: geometric ties (corner selection) need unfused
            # IEEE arithmetic on every host target
            cmd = [HIPCC] + COMMON + ["-ffp-contract=off", "-x", "c++", "-c", path, "-o", obj]
        out = _run(cmd)
        if verbose and out:
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    rocm_llvm = "/opt/rocm/llvm/lib"
    _run([HIPCC, "-shared", "-fopenmp", "-o", tmp] + objs +
         ["-Wl,-rpath," + rocm_llvm, "-L" + rocm_llvm])
    os.replace(tmp, LIB)
    shutil.rmtree(objdir, ignore_errors=True)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
