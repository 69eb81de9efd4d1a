"""Device readout (mvf_readout) against the oracle restatement of
node_solution / rx_data (itself pinned to the reference's solution.f90 and
receiver_data.f90 by tests/test_oracle_readout.py).

Same solution x on both sides (SciPy SuperLU on the system the device
assembles); bars: nodal E and H within 1e-11 of the largest value, receiver
E, H, Z within 1e-10 relative, rho_a within 1e-10 relative and phase within
1e-9 degrees (the same FP64 formulas in another summation order)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from movfem_amd import model as M
from movfem_amd.vfem import GpuVfem, MvfError
from oracle.oracle import Oracle
from tests.test_oracle_readout import CASES, air_layers, receivers_for

pytestmark = pytest.mark.gpu


def host_solution(g):
    ia, ja, a, b = g.global_vfem()
    ia, ja, a = g.remove_zeros(ia, ja, a)
    n = g.nne
    U = sp.coo_matrix((a, (ia - 1, ja - 1)), shape=(n, n)).tocsr()
    lu = spla.splu((U + sp.triu(U, 1).T).tocsc())
    b = b.reshape(2, n)
    return np.concatenate([lu.solve(b[d]) for d in range(2)])


def compare(m, x, es, hs, rx_out, xs, ys, ke_top):
    o = Oracle(m)
    o.set_frequency(0)
    oes, ohs = o.node_solution(x)
    for got, ref in ((es, oes), (hs, ohs)):
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()
    e, h, z, rho, phi = o.receivers(oes, ohs, ke_top, xs, ys)
    ge, gh, gz, grho, gphi = rx_out
    for got, ref in ((ge, e), (gh, h), (gz, z)):
        assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.allclose(grho, rho, rtol=1e-10, atol=0)
    assert np.abs(gphi - phi).max() <= 1e-9


@pytest.mark.parametrize("name", sorted(CASES))
def test_readout_matches_oracle(name, need_gpu):
    m = CASES[name]()
    print(f"readout {name}", flush=True)
    g = GpuVfem(m)
    g.set_frequency(0)
    x = host_solution(g)
    xs, ys = receivers_for(m)
    ke_top = m.nz - air_layers(m)
    out = g.readout(x, ke_top=ke_top, rx=xs, ry=ys)
    compare(m, x, out[0], out[1], out[2:], xs, ys, ke_top)


def test_edge_receivers(need_gpu):
    """Sites on and beyond the inner cells' edges inherit the previous
    receiver's element, as the reference's rx_data does (the oracle's rule is
    pinned to the reference build in test_oracle_readout.py)."""
    from tests.test_oracle_readout import edge_sites

    m = CASES["hex8_dirichlet"]()
    print("edge receivers", flush=True)
    g = GpuVfem(m)
    g.set_frequency(0)
    x = host_solution(g)
    xs, ys = edge_sites(m)
    ke_top = m.nz - air_layers(m)
    out = g.readout(x, ke_top=ke_top, rx=xs, ry=ys)
    compare(m, x, out[0], out[1], out[2:], xs, ys, ke_top)
    with pytest.raises(MvfError):
        g.readout(x, ke_top=ke_top, rx=xs[[3, 0]], ry=ys[[3, 0]])


def test_readout_after_device_solve(need_gpu):
    """Device-resident path: direct-preconditioned solve, readout of frequency 0
    without a host round trip of the solution."""
    m = CASES["hex8_gpml_aniso"]()
    g = GpuVfem(m)
    info = g.solve_frequencies([m.freqs[0]], method="gmres", precond="direct", tol=1e-10)
    assert info["converged"], info
    x = g.solution(2).ravel()
    xs, ys = receivers_for(m)
    ke_top = m.nz - air_layers(m)
    out = g.readout(None, f=0, ke_top=ke_top, rx=xs, ry=ys)
    compare(m, x, out[0], out[1], out[2:], xs, ys, ke_top)


@pytest.mark.parametrize("name", ["hex8_dirichlet", "hex20_gpml_aniso"])
def test_output_files_from_device(name, need_gpu):
    """SOLUTION.OUT, RECEIVER.OUT and SURFACE.OUT from the device readout
    (nodal Z, rho_a, phase and receiver heights included) against the same
    files from the oracle's fields: equal to print precision."""
    import pathlib
    import tempfile

    from movfem_amd import vfem
    from tests.outfiles import compare_files
    from tests.test_output_files import write_ours

    m = CASES[name]()
    print(f"output files {name}", flush=True)
    g = GpuVfem(m)
    g.set_frequency(0)
    x = host_solution(g)
    xs, ys = receivers_for(m)
    nza = air_layers(m)
    nzl_a, nzl_b = m.nzl_top, nza - m.nzl_top
    ke_top = m.nz - nza
    d = g.readout(x, ke_top=ke_top, rx=xs, ry=ys, full=True)
    o = Oracle(m)
    o.set_frequency(0)
    oes, ohs = o.node_solution(x)
    with tempfile.TemporaryDirectory() as td:
        td = pathlib.Path(td)
        ref = write_ours(m, oes, ohs, xs, ys, nzl_a, nzl_b, td)
        vfem.write_solution(td / "S", 1, d["esol"], d["hsol"], d["node_z"], d["node_rho"], d["node_phi"])
        vfem.write_receivers(td / "R", 1, 1, xs, ys, d["rx_elev"], d["rx_e"], d["rx_h"], d["rx_z"], d["rx_rho"],
                             d["rx_phi"])
        vfem.write_surface(td / "F", 1, 1, vfem.surface_nodes(m, nzl_a, nzl_b), d["esol"], d["hsol"], d["node_z"],
                           d["node_rho"], d["node_phi"])
        for ours, n in (("S", "SOLUTION.OUT"), ("R", "RECEIVER.OUT"), ("F", "SURFACE.OUT")):
            compare_files((td / ours).read_text(), ref[n], rel_tol=2.5e-5, abs_tol=2.5e-5, phase_branch=True)
