"""HIP path vs the CPU oracle on the same seeded inputs (through the C ABI).

Bars: DOF numbering, COO indices and the zero-removed pattern bit-exact;
element matrices and RHS within 1e-12 (normwise, relative); complex64 matrix
values within one float32 ulp of the oracle's."""
import numpy as np
import pytest

from movfem_amd import model as M
from movfem_amd.vfem import GpuVfem
from oracle.oracle import Oracle
from tests.parity_cases import cases

pytestmark = pytest.mark.gpu

CASES = cases()


def _rel(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s else 1.0)


@pytest.mark.parametrize("name", sorted(CASES))
def test_numbering_and_sizes(name, need_gpu):
    m = CASES[name]
    print(f"case {name}", flush=True)
    g = GpuVfem(m)
    o = Oracle(m)
    assert np.array_equal(g.gne(), o.gne())
    assert g.nne == o.nne
    assert g.nnze == o.nnze


def _element_matrices(name):
    m = CASES[name]
    print(f"case {name}", flush=True)
    g = GpuVfem(m)
    o = Oracle(m)
    g.set_frequency(0)
    o.set_frequency(0)
    K, Mm, b = g.integrate(0, m.ne)
    w32 = float(np.float32(m.omega(0)))
    me = m.edges_per_elem
    iu = np.triu_indices(me)
    ea = eb = 0.0
    e = 0
    for ie in range(1, m.nx + 1):
        for je in range(1, m.ny + 1):
            for ke in range(1, m.nz + 1):
                a_o, b_o = o.element(ie, je, ke)
                a_g = K[e] + 1j * (w32 * Mm[e])
                ea = max(ea, _rel(a_g, a_o[iu]))
                # alocal is symmetric in exact arithmetic
                assert _rel(a_o[iu], a_o.T[iu]) < 1e-12
                eb = max(eb, _rel(b[e], b_o))
                e += 1
    assert ea < 1e-12, ea
    assert eb < 1e-12, eb


@pytest.mark.parametrize("name", sorted(CASES))
def test_element_matrices(name, need_gpu):
    _element_matrices(name)


@pytest.mark.parametrize("name", ["hex8_gpml_aniso", "hex20_gpml_tensor_topo", "hex27_gpml_mu_topo"])
def test_element_matrices_valu_kernel(name, need_gpu, monkeypatch):
    """The one-workgroup-per-element VALU kernel (MVF_ELEM_KERNEL=valu, read when
    a handle is created) to the same bars."""
    monkeypatch.setenv("MVF_ELEM_KERNEL", "valu")
    _element_matrices(name)


@pytest.mark.parametrize("name", sorted(CASES))
def test_global_vfem(name, need_gpu):
    """Two frequencies in a row, as the reference's loop calls global_vfem."""
    m = CASES[name]
    print(f"case {name}", flush=True)
    g = GpuVfem(m)
    o = Oracle(m)
    for fi in range(2):
        g.set_frequency(fi)
        o.set_frequency(fi)
        ia, ja, a, b = g.global_vfem()
        ia_o, ja_o, a_o, b_o = o.global_vfem()
        assert np.array_equal(ia, ia_o) and np.array_equal(ja, ja_o)
        scale = np.abs(a_o).max()
        tol = 1.2e-7 * np.abs(a_o) + 1e-13 * scale
        assert np.all(np.abs(a - a_o) <= tol), np.abs(a - a_o).max()
        assert _rel(b, b_o) < 1e-12
        # find_zeros/rem_zeros: identical surviving pattern
        zi, zj, _ = g.remove_zeros(ia, ja, a)
        oi, oj, _ = Oracle.remove_zeros(ia_o, ja_o, a_o)
        assert np.array_equal(zi, oi) and np.array_equal(zj, oj)


@pytest.mark.parametrize("name", ["hex20_gpml_tensor_topo", "hex8_dirichlet_topo_mu"])
def test_global_vfem_chunks(name, monkeypatch, need_gpu):
    """global_vfem writes the coordinate form on the device in row chunks
    copied straight into the caller's arrays.  Chunks of 1 entry (one row
    each, since a row is never split), 97 entries and the default give the
    same bits, and a equals fl32(K) + i fl32(fl32(omega) M) of the device's
    operators on and above the diagonal and 0 below, with ia/ja the 1-based
    row and column of every pattern entry."""
    m = CASES[name]
    print(f"case {name}", flush=True)
    g = GpuVfem(m)
    g.set_frequency(0)
    w32 = np.float64(np.float32(m.omega(0)))

    def expected():
        rp, col, K, Mv = g.operators()
        rows = np.repeat(np.arange(g.nne), np.diff(rp))
        up = col >= rows
        re = np.where(up, K.astype(np.float32).astype(np.float64), 0.0)
        im = np.where(up, (w32 * Mv).astype(np.float32).astype(np.float64), 0.0)
        return rows + 1, col + 1, re + 1j * im

    got = []
    for ch in ("1", "97", None):
        if ch is None:
            monkeypatch.delenv("MVF_GV_CHUNK", raising=False)
        else:
            monkeypatch.setenv("MVF_GV_CHUNK", ch)
        ia, ja, a, _ = g.global_vfem()
        # call 0 uses the first-call GPML flags, later calls the carried ones
        # (the operators then change once); the operators read after each call
        # are the ones it formatted
        ia_e, ja_e, a_e = expected()
        assert np.array_equal(ia, ia_e) and np.array_equal(ja, ja_e)
        assert np.array_equal(a.view(np.float64), a_e.view(np.float64))
        got.append(a)
    assert np.array_equal(got[1].view(np.float64), got[2].view(np.float64))


@pytest.mark.parametrize("name", sorted(CASES))
def test_global_vfem_nz(name, monkeypatch, need_gpu):
    """mvf_global_vfem_nz equals mvf_global_vfem followed by mvf_remove_zeros
    (find_zeros/rem_zeros) bit for bit.  Two handles in lockstep over two
    frequencies (the second call switches to the carried GPML flags), with
    chunks of 1 entry (one row each), 97 entries and the default; then arrays
    one entry too short are refused."""
    import ctypes as C

    from movfem_amd import _lib as L

    m = CASES[name]
    print(f"case {name}", flush=True)
    g1, g2 = GpuVfem(m), GpuVfem(m)
    for fi, ch in ((0, "1"), (1, "97"), (1, None)):
        if ch is None:
            monkeypatch.delenv("MVF_GV_CHUNK", raising=False)
        else:
            monkeypatch.setenv("MVF_GV_CHUNK", ch)
        g1.set_frequency(fi)
        g2.set_frequency(fi)
        ia, ja, a, b = g1.global_vfem()
        zi, zj, za = g1.remove_zeros(ia, ja, a)
        ni, nj, na, nb = g2.global_vfem_nz()
        assert len(ni) <= g2.nnz_triu
        assert np.array_equal(ni, zi) and np.array_equal(nj, zj)
        assert np.array_equal(na.view(np.float64), za.view(np.float64))
        assert np.array_equal(nb.view(np.float64), b.view(np.float64))
    n = len(ni)
    ia = np.empty(n, np.int32)
    ja = np.empty(n, np.int32)
    a = np.empty(n, np.complex128)
    b = np.empty(2 * g2.nne, np.complex128)
    out = C.c_int64(-1)
    rc = L.lib().mvf_global_vfem_nz(g2.h, ia.ctypes.data_as(C.POINTER(C.c_int32)),
                                    ja.ctypes.data_as(C.POINTER(C.c_int32)),
                                    a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)),
                                    b.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), n - 1, g2.nne, 2,
                                    C.byref(out))
    assert rc == L.MVF_E_ARG and out.value == -1


def test_halfspace_block_global_vfem(need_gpu):
    """The 8-node Dirichlet half-space + block workload at full size."""
    m = M.halfspace_block()
    g = GpuVfem(m)
    o = Oracle(m)
    assert g.nne == o.nne == 21660
    g.set_frequency(0)
    o.set_frequency(0)
    ia, ja, a, b = g.global_vfem()
    ia_o, ja_o, a_o, b_o = o.global_vfem()
    assert np.array_equal(ia, ia_o) and np.array_equal(ja, ja_o)
    tol = 1.2e-7 * np.abs(a_o) + 1e-13 * np.abs(a_o).max()
    assert np.all(np.abs(a - a_o) <= tol)
    assert _rel(b, b_o) < 1e-12
