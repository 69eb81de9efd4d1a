"""GPU Krylov solve vs a direct FP64 solve of the same system.

The reference solves with MUMPS 4.10 (complex LDL^T of the complex64-rounded
upper triangle it assembles, MoVFEM_3DMT.f90:113-115), which is not available
here; SciPy's SuperLU (complex128) on exactly that matrix stands in for it.
Bars: FP64 true relative residual <= 1e-8 per polarization, and the solution
within 1e-6 (relative, 2-norm) of the direct solve."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from movfem_amd import model as M
from movfem_amd.vfem import GpuVfem

pytestmark = pytest.mark.gpu


def reference_system(g):
    """Matrix and RHS exactly as the reference hands them to MUMPS."""
    ia, ja, a, b = g.global_vfem()
    ia, ja, a = g.remove_zeros(ia, ja, a)
    n = g.nne
    U = sp.coo_matrix((a, (ia - 1, ja - 1)), shape=(n, n)).tocsr()
    A = (U + sp.triu(U, 1).T).tocsc()
    return A, b.reshape(2, n)


def _check(A, b, x, xref, tol_res=1e-8, tol_x=1e-6):
    for d in range(2):
        r = np.linalg.norm(b[d] - A @ x[d]) / np.linalg.norm(b[d])
        assert r <= tol_res * 1.0001, (d, r)
        e = np.linalg.norm(x[d] - xref[d]) / np.linalg.norm(xref[d])
        assert e <= tol_x, (d, e)


@pytest.mark.parametrize("method,precision", [("cocg", "fp64"), ("bicgstab", "fp64"), ("cocg", "mixed")])
def test_solve_small_dirichlet(method, precision, need_gpu):
    m = M.random_small(8, 4, 4, 4, seed=11, nz_air=1, freqs=(1.0,))
    print(f"solve {method}/{precision} nne", flush=True)
    g = GpuVfem(m)
    g.set_frequency(0)
    A, b = reference_system(g)
    lu = spla.splu(A)
    xref = np.stack([lu.solve(b[d]) for d in range(2)])
    x, info = g.solve(b.ravel(), method=method, precision=precision, tol=1e-8)
    assert info["converged"], info
    _check(A, b, x.reshape(2, -1), xref)


def test_solve_gpml_20node(need_gpu):
    m = M.random_small(20, 3, 3, 3, seed=12, dirichlet=False, nextd=1, anisotropic=True, freqs=(10.0,))
    g = GpuVfem(m)
    g.set_frequency(0)
    A, b = reference_system(g)
    lu = spla.splu(A)
    xref = np.stack([lu.solve(b[d]) for d in range(2)])
    x, info = g.solve(b.ravel(), tol=1e-9)
    assert info["converged"], info
    _check(A, b, x.reshape(2, -1), xref, tol_res=1e-9)


def test_batched_frequencies_match_sequential(need_gpu):
    """One batched sweep over 3 frequencies == three single-frequency solves."""
    freqs = (0.1, 1.0, 10.0)
    m = M.random_small(8, 4, 3, 4, seed=13, nz_air=1, freqs=freqs)
    gb = GpuVfem(m)
    info = gb.solve_frequencies(freqs, tol=1e-9)
    assert info["converged"], info
    xb = gb.solution(6)
    gs = GpuVfem(m)
    for i in range(3):
        gs.set_frequency(i)
        A, b = reference_system(gs)
        x, inf2 = gs.solve(b.ravel(), tol=1e-9)
        x = x.reshape(2, -1)
        bb = gb.rhs(6)[2 * i:2 * i + 2]
        assert np.allclose(bb, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())
        for d in range(2):
            assert np.linalg.norm(xb[2 * i + d] - x[d]) <= 1e-6 * np.linalg.norm(x[d])


def test_halfspace_block_solve(need_gpu):
    """The 8-node Dirichlet half-space + block workload (21,660 DOFs, 1 Hz).

    Three air layers above the first hold a gradient null space (see
    tests/gauge.py); the direct solve returns an arbitrary gauge there, so the
    1e-6 bar applies to the gauge-determined DOFs; the residual bar applies to
    every DOF."""
    from tests.gauge import determined_dofs

    m = M.halfspace_block()
    g = GpuVfem(m)
    g.set_frequency(0)
    A, b = reference_system(g)
    lu = spla.splu(A)
    xref = np.stack([lu.solve(b[d]) for d in range(2)])
    x, info = g.solve(b.ravel(), tol=1e-8)
    x = x.reshape(2, -1)
    print("halfspace_block", info["iterations"], info["seconds"], flush=True)
    assert info["converged"], info
    mask = determined_dofs(m, g.gne())
    assert 0.5 < mask.mean() < 1.0
    for d in range(2):
        r = np.linalg.norm(b[d] - A @ x[d]) / np.linalg.norm(b[d])
        assert r <= 1e-8 * 1.0001, (d, r)
        e = np.linalg.norm(x[d][mask] - xref[d][mask]) / np.linalg.norm(xref[d][mask])
        print(f"pol {d}: determined-DOF rel L2 {e:.2e}, full-vector {np.linalg.norm(x[d]-xref[d])/np.linalg.norm(xref[d]):.2e}")
        assert e <= 1e-6, (d, e)


def test_sixteen_frequencies_one_call(need_gpu):
    """The 27-node workload's 16 frequencies (1e-3 to 1e3 Hz) in one
    mvf_solve_frequencies call, which solves them in two batches of 8: every
    frequency to the bar on its own system, and a frequency of the second
    batch equal to its single-frequency solve.  Emulation off, so every
    frequency's system is consistent (the reference's update_sigma range
    leaves the second and later frequencies' air sources the null space
    cannot absorb)."""
    freqs = tuple(np.logspace(-3, 3, 16))
    m = M.random_small(8, 4, 4, 4, seed=14, nz_air=1, freqs=freqs)
    g = GpuVfem(m, emulate=0)
    print("16 frequencies: one call", flush=True)
    info = g.solve_frequencies(list(freqs), method="gmres", precond="direct", tol=1e-8)
    assert info["converged"], info
    assert len(info["true_resid"]) >= 32
    assert max(info["true_resid"][:32]) <= 1e-8 * 1.0001, info["true_resid"][:32]
    xb = g.solution(32)
    i = 12
    gs = GpuVfem(m, emulate=0)
    gs.set_frequency(i)
    A, b = reference_system(gs)
    assert np.allclose(g.rhs(32)[2 * i:2 * i + 2], b, rtol=1e-12, atol=1e-12 * np.abs(b).max())
    x, _ = gs.solve(b.ravel(), method="gmres", precond="direct", tol=1e-8)
    x = x.reshape(2, -1)
    xref = np.stack([spla.splu(A).solve(b[d]) for d in range(2)])
    _check(A, b, xb[2 * i:2 * i + 2], xref)
    _check(A, b, x, xref)
