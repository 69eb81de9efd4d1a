"""mvf_solve_noalloc (no device allocation; Krylov buffers in a static device
array) against mvf_solve, through the C ABI: two calls in a row on one GPU,
the second on a smaller model, each for COCG and BiCGStab.  Both entry points
run the same kernels in the same order, so the solutions agree to the last
bit; the bar asserted is the FP64 true relative residual <= 1e-8 and
agreement within 1e-12 (relative, 2-norm)."""
import ctypes as C

import numpy as np
import pytest

from movfem_amd import _lib as L
from movfem_amd import model as M
from movfem_amd.vfem import GpuVfem, MvfError

pytestmark = pytest.mark.gpu


def _system(g):
    import scipy.sparse as sp

    ia, ja, a, b = g.global_vfem()
    ia, ja, a = g.remove_zeros(ia, ja, a)
    n = g.nne
    U = sp.coo_matrix((a, (ia - 1, ja - 1)), shape=(n, n)).tocsr()
    return (U + sp.triu(U, 1).T).tocsr(), b.reshape(2, n)


@pytest.mark.parametrize("method", ["cocg", "bicgstab"])
def test_noalloc_matches_solve(method, need_gpu):
    models = [M.halfspace_block(), M.random_small(8, 4, 4, 4, seed=11, nz_air=1, freqs=(1.0,))]
    sizes = []
    for m in models:
        print(f"noalloc {method} nne", flush=True)
        g = GpuVfem(m)
        g.set_frequency(0)
        A, b = _system(g)
        sizes.append(g.nne)
        x1, i1 = g.solve(b.ravel(), method=method, tol=1e-8)
        x2, i2 = g.solve_noalloc(b.ravel(), method=method, tol=1e-8)
        print(f"{method} nne={g.nne}: iterations {i1['iterations']} / {i2['iterations']}, bit-identical "
              f"{np.array_equal(x1, x2)}", flush=True)
        assert i2["converged"] and i2["iterations"] == i1["iterations"]
        x2 = x2.reshape(2, -1)
        x1 = x1.reshape(2, -1)
        for d in range(2):
            r = np.linalg.norm(b[d] - A @ x2[d]) / np.linalg.norm(b[d])
            assert r <= 1e-8 * 1.0001, (d, r)
            assert np.linalg.norm(x2[d] - x1[d]) <= 1e-12 * np.linalg.norm(x1[d])
        g.close()
    assert sizes[1] < sizes[0]


def test_noalloc_limits(need_gpu):
    m = M.random_small(8, 3, 3, 3, seed=12, nz_air=1, freqs=(1.0,))
    g = GpuVfem(m)
    g.set_frequency(0)
    b = np.ones(2 * g.nne, np.complex128)
    with pytest.raises(MvfError) as e:
        g.solve_noalloc(b, method="gmres", precond="direct")
    assert e.value.code == L.MVF_E_UNSUPPORTED
    big = M.halfspace_block(nx=52, ny=52, nz_earth=36, nz_air=4)
    gb = GpuVfem(big)
    gb.set_frequency(0)
    assert gb.nne > 299593
    with pytest.raises(MvfError) as e:
        gb.solve_noalloc(np.ones(2 * gb.nne, np.complex128), method="cocg")
    assert e.value.code == L.MVF_E_ARG and "scratch" in str(e.value)
