"""One-element-wide grids, the smallest inputs the reference's loops accept.

8-node 1 x 1 x 2 with Dirichlet boundaries has no unknowns at all (every edge
lies on the boundary): the oracle numbers none, and the library refuses the
model at mvf_init with MVF_E_ARG instead of building an empty system.  The
other grids (2 to 68 unknowns) go through numbering, element matrices and
global_vfem against the oracle to the parity bars of test_gpu_parity.py."""
import numpy as np
import pytest

from movfem_amd import _lib as L
from movfem_amd import model as M
from movfem_amd.vfem import GpuVfem, MvfError
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

GRIDS = [(8, (2, 1, 3)), (8, (1, 3, 2)), (20, (1, 1, 2)), (20, (2, 1, 3)), (27, (1, 1, 2)), (27, (1, 3, 2))]


def _rel(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s else 1.0)


def test_no_unknowns_is_refused(need_gpu):
    m = M.random_small(8, 1, 1, 2, seed=41, nz_air=1)
    assert Oracle(m).nne == 0
    print("tiny: 8-node 1x1x2 Dirichlet, no unknowns", flush=True)
    with pytest.raises(MvfError) as e:
        GpuVfem(m)
    assert e.value.code == L.MVF_E_ARG


@pytest.mark.parametrize("nodes,dims", GRIDS)
def test_tiny_grid_parity(nodes, dims, need_gpu):
    m = M.random_small(nodes, *dims, seed=41, nz_air=1)
    print(f"tiny: {nodes}-node {dims}", flush=True)
    g = GpuVfem(m)
    o = Oracle(m)
    assert np.array_equal(g.gne(), o.gne())
    assert g.nne == o.nne and g.nnze == o.nnze
    g.set_frequency(0)
    o.set_frequency(0)
    K, Mm, b = g.integrate(0, m.ne)
    w32 = float(np.float32(m.omega(0)))
    iu = np.triu_indices(m.edges_per_elem)
    e = 0
    for ie in range(1, m.nx + 1):
        for je in range(1, m.ny + 1):
            for ke in range(1, m.nz + 1):
                a_o, b_o = o.element(ie, je, ke)
                assert _rel(K[e] + 1j * (w32 * Mm[e]), a_o[iu]) < 1e-12
                assert _rel(b[e], b_o) < 1e-12
                e += 1
    g2, o2 = GpuVfem(m), Oracle(m)
    g2.set_frequency(0)
    o2.set_frequency(0)
    ia, ja, a, bb = g2.global_vfem()
    ia_o, ja_o, a_o, b_o = o2.global_vfem()
    assert np.array_equal(ia, ia_o) and np.array_equal(ja, ja_o)
    assert np.all(np.abs(a - a_o) <= 1.2e-7 * np.abs(a_o) + 1e-13 * np.abs(a_o).max())
    assert _rel(bb, b_o) < 1e-12


@pytest.mark.parametrize("nodes,dims", [(20, (2, 1, 3)), (27, (1, 1, 2)), (27, (1, 3, 2))])
def test_tiny_grid_direct_solve(nodes, dims, need_gpu):
    """A nested-dissection tree of one front (no update DOFs) or very few
    through the factorization and GMRES, judged on the reference's matrix."""
    import scipy.sparse as sp

    m = M.random_small(nodes, *dims, seed=41, nz_air=1)
    print(f"tiny solve: {nodes}-node {dims}", flush=True)
    g = GpuVfem(m)
    g.set_frequency(0)
    ia, ja, a, b = g.global_vfem()
    ia, ja, a = g.remove_zeros(ia, ja, a)
    n = g.nne
    U = sp.coo_matrix((a, (ia - 1, ja - 1)), shape=(n, n)).tocsr()
    A = (U + sp.triu(U, 1).T).tocsc()
    b = b.reshape(2, n)
    x, info = g.solve(b.ravel(), method="gmres", precond="direct", tol=1e-8)
    x = x.reshape(2, -1)
    assert info["converged"], info
    for d in range(2):
        r = np.linalg.norm(b[d] - A @ x[d]) / np.linalg.norm(b[d])
        assert r <= 1e-8 * 1.0001, (d, r)


def test_element_count_limit(need_gpu):
    """The element index is decomposed in 32-bit arithmetic, so mvf_init
    refuses more than 2^31 - 1 elements (2048 x 2048 x 1024 here) with
    MVF_E_ARG before it reads the lattice or model arrays (which are
    one-element stand-ins here) or allocates anything for them."""
    import ctypes as C

    lib = L.lib()
    one = np.zeros(2)
    dp = one.ctypes.data_as(C.POINTER(C.c_double))
    grid = L.Grid(2048, 2048, 1024, 8, dp, dp, dp)
    model = L.Model(dp, dp)
    lz = np.zeros(2)
    bc = L.BC(1, 1, 0, 1, 1, 0, 1.0, 1.0, 2.0, 0.0, 0.01, 1, 0, lz.ctypes.data_as(C.POINTER(C.c_double)),
              lz.ctypes.data_as(C.POINTER(C.c_double)))
    opts = L.Options(L.MVF_EMU_ALL, 0, 0, 0)
    h = C.c_void_p()
    print("element count limit: 2048 x 2048 x 1024", flush=True)
    rc = lib.mvf_init(C.byref(grid), C.byref(model), C.byref(bc), C.byref(opts), C.byref(h))
    msg = lib.mvf_last_error(h).decode() if h.value else ""
    if h.value:
        lib.mvf_free(h)
    assert rc == L.MVF_E_ARG, (rc, msg)
    assert "2^31" in msg, msg


def test_frequency_count_limits(need_gpu):
    """mvf_solve_frequencies takes 1..32 frequencies per call (the 27-node
    workload's 16 fit one call); none or 33 are refused with MVF_E_ARG."""
    m = M.random_small(27, 1, 1, 2, seed=41, nz_air=1)
    g = GpuVfem(m)
    for freqs in ([], list(np.logspace(-3, 3, 33))):
        print(f"frequency count {len(freqs)}", flush=True)
        with pytest.raises(MvfError) as e:
            g.solve_frequencies(freqs, method="gmres", precond="direct", tol=1e-8)
        assert e.value.code == L.MVF_E_ARG
    info = g.solve_frequencies(list(np.logspace(-3, 3, 32)), method="gmres", precond="direct", tol=1e-8)
    assert info["converged"]
