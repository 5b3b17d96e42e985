"""Builds libed_hip.so (HIP kernels + C-ABI) for gfx950 and the Fortran H_class shim.

    python -m ed_solver_amd.build        # or __graft_entry__.build()

hipcc cross-compiles for gfx950 without a GPU.  The Fortran module and its demo
driver are compiled with amdflang when it is present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libed_hip.so")
SOURCES = ["ed_runtime.hip", "ed_tables.hip", "ed_engine.hip", "ed_green.hip", "ed_batch.hip", "ed_block.hip", "ed_arpack.hip", "ed_grad.hip", "ed_noalloc.hip", "ed_tridiag.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("ED_FLANG", "/opt/rocm/llvm/bin/amdflang")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unused-result",
    # packed FP32 ops stay disabled (hand-written or compiler-made v_pk_*_f32 are avoided)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # host code (the band and tridiagonal eigen-solvers of the Ritz tests) with AVX2 and FMA:
    # every x86-64 host an MI355X sits in has them; the block Lanczos Ritz test runs 1.7x faster
    "-Xarch_host", "-mavx2", "-Xarch_host", "-mfma",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_hip() -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, *srcs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_fortran() -> str | None:
    """Compile the ISO_C_BINDING H_class module and its demo driver (needs amdflang)."""
    fdir = os.path.join(PKG, "fortran")
    if not shutil.which(FLANG) and not os.path.exists(FLANG):
        return None
    bindir = os.path.join(fdir, "bin")
    os.makedirs(bindir, exist_ok=True)
    exe = os.path.join(bindir, "ed_lanczos_demo")
    _run([FLANG, "-O2", "-J", bindir, os.path.join(fdir, "ed_hip_mod.f90"), os.path.join(fdir, "h_class_hip.f90"),
          os.path.join(fdir, "ed_lanczos_demo.f90"), "-L", PKG, "-led_hip", f"-Wl,-rpath,{PKG}", "-o", exe], cwd=fdir)
    return exe


def build_all() -> None:
    build_hip()
    if os.path.exists(os.path.join(PKG, "fortran", "h_class_hip.f90")):
        build_fortran()


if __name__ == "__main__":
    build_all()
    print(LIB)
