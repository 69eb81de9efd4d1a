"""PROBLEM.OUT's fields on the device (k_problem through mvf_problem_fields)
against the oracle, whose values equal the reference build's exactly
(tests/test_output_files.py): the primary E and H and the source term at
every lattice node, from the last element holding it, at two frequencies
(the second after update_sigma).  Bar: 1e-13 of each array's largest value
per polarization; the file written from the device values carries the same
numbers to that bar."""
import pathlib
import tempfile

import numpy as np
import pytest

from movfem_amd import vfem
from movfem_amd.vfem import GpuVfem
from oracle.oracle import Oracle
from tests.outfiles import parse_problem
from tests.parity_cases import cases

pytestmark = pytest.mark.gpu
NAMES = ["hex8_dirichlet_topo_mu", "hex8_gpml_aniso", "hex20_gpml_tensor_topo", "hex27_gpml_mu_topo",
         "hex27_dirichlet_tensor"]


def _rel(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s else 1.0)


@pytest.mark.parametrize("name", NAMES)
def test_problem_fields_device(name, need_gpu):
    m = cases()[name]
    print(f"problem fields {name}", flush=True)
    g = GpuVfem(m)
    o = Oracle(m)
    for fi in range(2):
        g.set_frequency(fi)
        o.set_frequency(fi)
        dev, ref = g.problem_fields(), o.problem_fields()
        for what, a, b in zip(("ep", "hp", "src"), dev, ref):
            for d in range(2):
                assert _rel(a[d], b[d]) <= 1e-13, (what, d, _rel(a[d], b[d]))
    with tempfile.TemporaryDirectory() as td:
        p = pathlib.Path(td) / "PROBLEM.OUT"
        vfem.write_problem(p, 2, *dev)
        blocks = parse_problem(p.read_text())
    for d, (head, arr) in enumerate(blocks):
        assert arr.shape == (m.npt, 10)
        for k, b in enumerate(ref):
            assert _rel(arr[:, 1 + 3 * k:4 + 3 * k], b[d]) <= 1e-13
