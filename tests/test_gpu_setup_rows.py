"""The per-row pattern build and assembly (large models) against the pair sort:
the same CSR pattern, DOF counts and FP64 values bit for bit on the parity
cases (8-, 20- and 27-node, Dirichlet and GPML, topography, tensor sigma, mu
contrast), and the same FP32-integrated values.  Both paths sum every entry
over its elements in element order, the reference's order
(MoVFEM_3DMT.f90:193-211, global_assembly.f90:43-60)."""
import os

import numpy as np
import pytest

from movfem_amd.vfem import GpuVfem
from tests.parity_cases import cases

pytestmark = pytest.mark.gpu

CASES = cases()


def _build(m, path):
    old = os.environ.get("MVF_SETUP")
    os.environ["MVF_SETUP"] = path
    try:
        return GpuVfem(m)
    finally:
        if old is None:
            del os.environ["MVF_SETUP"]
        else:
            os.environ["MVF_SETUP"] = old


@pytest.mark.parametrize("name", sorted(CASES))
def test_rows_equal_pairs(name, need_gpu):
    m = CASES[name]
    print(f"setup paths {name}", flush=True)
    gp, gr = _build(m, "pairs"), _build(m, "rows")
    assert (gp.nne, gp.nnze, gp.nnz_triu) == (gr.nne, gr.nnze, gr.nnz_triu)
    for a, b in zip(gp.operators(), gr.operators()):
        assert np.array_equal(a, b)
    gp.assemble("fp32")
    gr.assemble("fp32")
    for a, b in zip(gp.operators()[2:], gr.operators()[2:]):
        assert np.array_equal(a, b)
