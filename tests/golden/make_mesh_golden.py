"""Writes the mesh-generation golden vectors from the reference build
(oracle/_ref: the reference's own read_input.f90, toms660.f90 and
geometry.f90, compiled by oracle/Makefile).

Run in the container that holds the reference sources:
    make -C oracle ref && python -m tests.golden.make_mesh_golden [name ...]

For the reference's example input (sample_3DEM_FEM.INP with the frequencies of
sample_PARAM.INP) and for every synthetic case of tests/inputfiles.py, the
reference runs input() on the file and grid_3d on what it read.  Each
mesh_<name>.npz stores the reader's arrays and the lattice: small arrays as
they are, large ones (g_zp, g_sigma, g_mu and the interface and model arrays)
as SHA-256 digests of their bytes plus a strided sample for diagnosis.
"""
from __future__ import annotations

import hashlib
import pathlib
import sys

import numpy as np

from oracle.refrun import run_reference_grid
from tests.inputfiles import CASES, case_files

HERE = pathlib.Path(__file__).resolve().parent
SMALL = ("in_isz", "in_rsz", "in_xsr", "in_ysr", "in_zsr", "in_nsp", "in_xm", "in_ym", "in_ijsigma", "in_ijmu",
         "g_isz", "g_rsz", "g_xp", "g_yp", "g_nzl", "g_xsr", "g_ysr", "g_zsr")
LARGE = ("in_xto", "in_yto", "in_zto", "in_zm", "in_sigma", "in_mu", "g_zp", "g_sigma", "g_mu")


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sample(a, k=199):
    a = np.ascontiguousarray(a).ravel()
    step = max(1, len(a) // k)
    return a[::step].copy()


def sample_files():
    text = (HERE / "sample_3DEM_FEM.INP").read_text()
    return text, [0.1]  # sample_PARAM.INP: one frequency, 1.e-1 Hz


def write(name):
    text, freqs = sample_files() if name == "sample" else case_files(name)
    r = run_reference_grid(text, freqs, timeout=1200)
    out = {"freqs": np.asarray(freqs, np.float64)}
    for k in SMALL:
        out[k] = r[k]
    for k in LARGE:
        out[k + "_sha256"] = np.asarray(digest(r[k]))
        out[k + "_sample"] = sample(r[k])
    np.savez_compressed(HERE / f"mesh_{name}.npz", **out)
    print(name, r["g_isz"], r["g_nzl"], flush=True)


def main(names):
    for n in names or ["sample", *CASES]:
        write(n)


if __name__ == "__main__":
    main(sys.argv[1:])
