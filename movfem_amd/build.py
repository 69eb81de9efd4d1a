"""Builds libmovfem_gpu.so (gfx950) in-tree with hipcc.

Run as `python -m movfem_amd.build`.  The library has no torch dependency: its
ABI is plain C (include/movfem_gpu.h).
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "libmovfem_gpu.so"
SRC = PKG / "csrc" / "mvf_lib.hip"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # never emit packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> pathlib.Path:
    deps = sorted((PKG / "csrc").glob("*")) + [ROOT / "include" / "movfem_gpu.h"]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(ROOT / "include"), str(SRC), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout[-4000:] + r.stderr[-8000:])
    os.replace(tmp, LIB)
    _ = deps
    build_fortran(verbose)
    return LIB


FDIR = PKG / "fortran"
FDRIVER = FDIR / "movfem_driver"
FRUNCASE = FDIR / "run_case"


def build_fortran(verbose: bool = False) -> pathlib.Path | None:
    """The ISO_C_BINDING module, the Fortran driver and run_case (needs amdflang)."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if not fc:
        return None
    mod = FDIR / "_mod"
    mod.mkdir(exist_ok=True)
    for prog, out in (("driver.f90", FDRIVER), ("run_case.f90", FRUNCASE)):
        cmd = [fc, "-O2", "-module-dir", str(mod), str(FDIR / "movfem_gpu_mod.f90"), str(FDIR / prog),
               "-L", str(PKG), "-lmovfem_gpu", "-Wl,-rpath,$ORIGIN/..", "-o", str(out)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("amdflang failed:\n" + r.stdout[-2000:] + r.stderr[-4000:])
    return FDRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
