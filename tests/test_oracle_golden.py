"""The CPU oracle against golden vectors written by the reference's own
modules (tests/golden/make_golden.py drives oracle/_ref).  Indices bit-exact,
element matrices and RHS to 1e-13, complex64 values to one float32 ulp."""
import pathlib

import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.golden.make_golden import load_model

# model fixtures (out_*.npz and ld_reals.npz hold output files and printed
# reals, tests/test_output_files.py; mesh_*.npz the lattice generator's,
# tests/test_mesh.py)
GOLD = sorted(p for p in (pathlib.Path(__file__).parent / "golden").glob("*.npz")
              if not p.stem.startswith(("out_", "mesh_", "ld_")))


def _rel(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s else 1.0)


@pytest.mark.parametrize("path", GOLD, ids=[p.stem for p in GOLD])
def test_golden(path):
    z = np.load(path)
    m = load_model(z)
    o = Oracle(m)
    assert np.array_equal(o.gne(), z["gne"])
    assert np.array_equal(o.nzindx(), z["nzindx"])
    assert o.nne == int(z["nne"]) and o.nnze == int(z["nnze"])
    o.set_frequency(0)
    e = 0
    for ie in range(1, m.nx + 1):
        for je in range(1, m.ny + 1):
            for ke in range(1, m.nz + 1):
                a, b = o.element(ie, je, ke)
                assert _rel(a, z["aloc"][e]) < 1e-13
                assert _rel(b, z["bloc"][e]) < 1e-13
                e += 1
    # global_vfem for frequencies 1 and 2; the reference ran the element sweep
    # first, which advanced its GPML flag state exactly as the sweep above did
    for fi in (1, 2):
        o.set_frequency(fi - 1)
        ia, ja, a, b = o.global_vfem()
        assert np.array_equal(ia, z[f"ia_{fi}"]) and np.array_equal(ja, z[f"ja_{fi}"])
        ar = z[f"a_{fi}"]
        assert np.all(np.abs(a - ar) <= 1.2e-7 * np.abs(ar) + 1e-14 * np.abs(ar).max())
        assert _rel(b, z[f"b_{fi}"]) < 1e-13


@pytest.mark.parametrize("path", GOLD, ids=[p.stem for p in GOLD])
def test_golden_update_sigma(path):
    """update_sigma's linear-index range (geometry.f90:144-153) at frequency 2."""
    z = np.load(path)
    m = load_model(z)
    o = Oracle(m)
    o.set_frequency(0)
    o.set_frequency(1)
    assert np.array_equal(o.sigma(), z["sigma_2"])


BD = [p for p in GOLD if "h_sigma" in np.load(p).files]


@pytest.mark.parametrize("path", BD, ids=[p.stem for p in BD])
def test_golden_dirichlet_readout(path):
    """node_solution with the inimod 2/3 Dirichlet values on the boundary edges
    (solution.f90:327-329) at frequency 2, where the reference's run left it."""
    z = np.load(path)
    m = load_model(z)
    o = Oracle(m)
    o.set_frequency(0)
    o.set_frequency(1)
    e, h = o.node_solution(z["readout_x"])
    assert _rel(e, z["esol"]) < 1e-13
    assert _rel(h, z["hsol"]) < 1e-13
