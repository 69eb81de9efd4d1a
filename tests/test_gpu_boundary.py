"""Dirichlet inimod 2/3 on the device against the oracle (through the C ABI).

The boundary values (f_boundary with the half-space and layered-earth primary
fields, boundary_conds.f90:188-598) enter the system through the element RHS
lift (MoVFEM_3DMT.f90:251-261) and the readout of the Dirichlet edges
(solution.f90:327-329); the oracle's restatement of both is pinned to the
reference build (tests/test_oracle_boundary.py).  Bars as for the rest of the
path: RHS within 1e-12 (normwise, relative), nodal E and H within 1e-11 of the
largest value, receivers within 1e-10; the device direct solve reaches a true
residual of 1e-8 on the reference's complex64 matrix and agrees with SuperLU
to 1e-6 on the gauge-determined DOFs."""
import dataclasses

import numpy as np
import pytest

from movfem_amd.vfem import GpuVfem
from oracle.oracle import Oracle
from tests.gauge import determined_dofs
from tests.parity_cases import cases
from tests.test_gpu_direct import reference_system
from tests.test_gpu_readout import compare, host_solution
from tests.test_oracle_readout import CASES as RCASES
from tests.test_oracle_readout import air_layers, receivers_for

pytestmark = pytest.mark.gpu

BD = {"halfspace": dict(inimod=2, h_sigma=0.02),
      "layers3": dict(inimod=3, l_sigma=(0.05, 0.002, 0.3), l_dz=(0.4, 1.1)),
      "layer1": dict(inimod=3, l_sigma=(0.03,), l_dz=())}
NAMES = sorted(n for n, m in cases().items() if m.dirichlet)


def model(name, bd):
    return dataclasses.replace(cases()[name], **BD[bd])


def _rel(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s else 1.0)


@pytest.mark.parametrize("bd", sorted(BD))
@pytest.mark.parametrize("name", NAMES)
def test_rhs_lift(name, bd, need_gpu):
    m = model(name, bd)
    print(f"lift {name} {bd}", flush=True)
    g = GpuVfem(m)
    o = Oracle(m)
    for fi in range(2):
        g.set_frequency(fi)
        o.set_frequency(fi)
        ia, ja, a, b = g.global_vfem()
        ia_o, ja_o, a_o, b_o = o.global_vfem()
        assert np.array_equal(ia, ia_o) and np.array_equal(ja, ja_o)
        assert _rel(b, b_o) < 1e-12, _rel(b, b_o)


RNAMES = sorted(n for n in RCASES if "dirichlet" in n)


@pytest.mark.parametrize("bd", sorted(BD))
@pytest.mark.parametrize("name", RNAMES)
def test_readout_dirichlet_edges(name, bd, need_gpu):
    m = dataclasses.replace(RCASES[name](), **BD[bd])
    print(f"readout {name} {bd}", flush=True)
    g = GpuVfem(m)
    g.set_frequency(0)
    x = host_solution(g)
    xs, ys = receivers_for(m)
    ke_top = m.nz - air_layers(m)
    out = g.readout(x, ke_top=ke_top, rx=xs, ry=ys)
    compare(m, x, out[0], out[1], out[2:], xs, ys, ke_top)


@pytest.mark.parametrize("bd", sorted(BD))
def test_device_solve(bd, need_gpu):
    """Device-built RHS (with the lift) solved by the device factorization,
    against SuperLU on the system global_vfem hands over."""
    m = model("hex20_dirichlet", bd)
    print(f"solve {bd}", flush=True)
    g = GpuVfem(m)
    g.set_frequency(0)
    A, b = reference_system(g)
    info = g.solve_frequencies([m.freqs[0]], method="gmres", precond="direct", tol=1e-10)
    assert info["converged"], info
    assert np.abs(g.rhs(2) - b).max() <= 1e-12 * np.abs(b).max()
    x = g.solution(2)
    mask = determined_dofs(m, g.gne())
    from scipy.sparse.linalg import splu

    lu = splu(A)
    for d in range(2):
        r = np.linalg.norm(b[d] - A @ x[d]) / np.linalg.norm(b[d])
        assert r <= 1e-8, (d, r)
        xr = lu.solve(b[d])
        assert np.linalg.norm(x[d][mask] - xr[mask]) <= 1e-6 * np.linalg.norm(xr[mask])
