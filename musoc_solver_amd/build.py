"""Build the native parts of the package in-tree.

  libsevp_hip.so     HIP kernels + C ABI (include/sevp.h), gfx950
  libsevp_fortran.so ISO_C_BINDING module sevp_hip and the drop-in
                     REPBIR / REP1 / REP / PRE / PRE1 / MATINV wrappers,
                     linked against libsevp_hip.so

Run ``python -m musoc_solver_amd.build`` (or __graft_entry__.build()).
Both libraries land in musoc_solver_amd/lib/.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libsevp_hip.so")
FLIB = os.path.join(LIBDIR, "libsevp_fortran.so")
HIP_SOURCES = ["sevp_bir.hip", "sevp_full.hip", "sevp_matinv.hip", "sevp_mixed.hip", "sevp_fused.hip", "sevp_caller.hip",
               "sevp_krylov.hip", "sevp_capi.hip", "sevp_evp.cpp"]
ARCH = "gfx950"

# -ffp-contract=off: every multiply and add of the parity path is separately
# rounded, as in the reference.  Packed FP32 instructions are disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _flang() -> str | None:
    for c in (shutil.which("flang"), "/opt/rocm/llvm/bin/flang", "/opt/rocm/bin/amdflang"):
        if c and os.path.exists(c):
            return c
    return None


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]}")
    return r


def build_hip(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    hipcc = _hipcc()
    for src in HIP_SOURCES:
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    fc = _flang()
    src = os.path.join(FSRC, "sevp_hip.f90")
    if fc is None or not os.path.exists(src):
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = FLIB + ".tmp"
    cmd = [fc, "-O2", "-fPIC", "-shared", "-J", LIBDIR, src, "-o", tmp,
           f"-L{LIBDIR}", "-lsevp_hip", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, FLIB)
    return FLIB


DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "dropin_driver.f90")
DRIVER_EXE = os.path.join(ROOT, "tests", "fortran", "dropin_driver.exe")


def build_fortran_driver(verbose: bool = False) -> str | None:
    """The test program that calls the drop-in procedures as the reference's
    callers do (tests/test_gpu_fortran.py runs it)."""
    fc = _flang()
    if fc is None or not os.path.exists(DRIVER_SRC) or not os.path.exists(FLIB):
        return None
    tmp = DRIVER_EXE + ".tmp"
    rpath = os.path.relpath(LIBDIR, os.path.dirname(DRIVER_EXE))
    cmd = [fc, "-O2", f"-I{LIBDIR}", DRIVER_SRC, "-o", tmp, f"-L{LIBDIR}", "-lsevp_fortran",
           "-lsevp_hip", f"-Wl,-rpath,$ORIGIN/{rpath}"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd, cwd=os.path.dirname(DRIVER_EXE))
    os.replace(tmp, DRIVER_EXE)
    return DRIVER_EXE


def build(verbose: bool = False) -> None:
    build_hip(verbose)
    build_fortran(verbose)
    build_fortran_driver(verbose)


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
