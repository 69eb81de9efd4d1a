"""Writes tests/golden/ld_reals.npz: doubles and the text the reference's
compiler (amdflang, the compiler oracle/Makefile builds the reference with)
prints for them with list-directed output, one value per record.

    python -m tests.golden.make_ld_golden

The values mix magnitudes from 1e-300 to 1e300, short decimals, values that
need 17 digits, powers of two and ten and their neighbours, and subnormals."""
from __future__ import annotations

import pathlib
import shutil
import subprocess
import tempfile

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent

PROGRAM = """program ld
  use iso_fortran_env, only: real64
  implicit none
  integer :: n, i, u
  real(real64), allocatable :: x(:)
  open(newunit=u, file='x.bin', access='stream', form='unformatted', status='old')
  read(u) n
  allocate(x(n))
  read(u) x
  close(u)
  do i = 1, n
    write(*, *) x(i)
  end do
end program ld
"""


def values(seed=11):
    rng = np.random.Generator(np.random.PCG64(seed))
    v = [10.0 ** rng.uniform(-300, 300, 1500), 10.0 ** rng.uniform(-8, 20, 2500),
         np.array([round(a, int(d)) for a, d in zip(rng.uniform(-1e4, 1e4, 500), rng.integers(0, 6, 500))]),
         rng.uniform(0, 1, 500) * 10.0 ** rng.integers(-5, 16, 500)]
    p = 2.0 ** np.arange(-60, 61)
    t = 10.0 ** np.arange(-20, 21)
    edges = np.concatenate([p, np.nextafter(p, 0), np.nextafter(p, np.inf), t, np.nextafter(t, 0),
                            np.nextafter(t, np.inf), [0.1, 0.2, 0.3, 1 / 3, 2 / 3, 5e-324, 2.2e-308, 1.7e308]])
    x = np.concatenate(v + [edges])
    sign = rng.choice([-1.0, 1.0], len(x))
    return x * sign


def main():
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    x = values()
    with tempfile.TemporaryDirectory() as td:
        td = pathlib.Path(td)
        (td / "ld.f90").write_text(PROGRAM)
        subprocess.run([fc, "-O1", "ld.f90", "-o", "ld"], cwd=td, check=True, capture_output=True)
        with open(td / "x.bin", "wb") as f:
            np.array([len(x)], np.int32).tofile(f)
            x.astype(np.float64).tofile(f)
        out = subprocess.run(["./ld"], cwd=td, check=True, capture_output=True, text=True).stdout
    text = [ln.strip() for ln in out.splitlines()]
    assert len(text) == len(x)
    np.savez_compressed(HERE / "ld_reals.npz", x=x, text=np.array(text))
    print(len(x), "values")


if __name__ == "__main__":
    main()
