"""Mixed precision on the device: FP32 element integration, with its error
against the FP64 path stated per result.

FP32 integration (mvf_assemble precision 1) evaluates int_elem_params and
alocal's sums (integration.f90:60-106) in float32 and assembles K and M from
those element sums.  Measured on the device (8-node config 1, an 8-node GPML
anisotropic topography case of 45 elements, which ends the four-elements-per-
wave kernel on a ragged group, and 20- and 27-node GPML cases): every
assembled entry within 1.4e-4 * sqrt(|X_ii X_jj|) of the FP64 one (X = K or
M), Frobenius-relative within 4e-7.  Bars: 2e-4 and 1e-6.  The largest
normalized errors (1.4e-4; the 8- and 27-node cases stay below 3.1e-6) sit on
20-node mass entries whose element sums cancel (serendipity functions change
sign at the Gauss points); the normalization by the diagonal entries is then
small against the terms summed.

The solution of the FP32-integrated operator, rounded to complex64 as the
reference rounds what it hands MUMPS (FP64 direct-preconditioned GMRES to an
FP64 true residual <= 1e-8 on that matrix), against SciPy SuperLU on the
FP64-integrated reference matrix (config 1).  Measured: secondary E 1.6e-4
(relative L2, gauge-determined DOFs), rho_a 5.4e-5, phase 1.8e-3 degrees, and
120 GMRES iterations against 2 for the FP64-integrated matrix (7e-11, 1.5e-11,
5e-10 degrees): float32 sums leave the air's gradient null space inexact, so
near-null modes the factorization of the shifted matrix does not resolve
remain for GMRES.  Bars: 5e-4, 0.1 %, 0.05 degrees; the FP64 row 1e-6.

FP32 inner Krylov iterations inside FP64 refinement (mixed COCG) converge on
small Dirichlet models (tests/test_gpu_solve.py) but stagnate on config 1:
its air null space and the complex64 operator rounding put kappa * eps32
above one, so the FP32 correction equation cannot be reduced by the inner
tolerance.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from movfem_amd import _lib as L
from movfem_amd import model as M
from movfem_amd.vfem import GpuVfem, MvfError
from oracle.oracle import Oracle
from tests.gauge import determined_dofs
from tests.test_gpu_configs import config1_receivers

pytestmark = pytest.mark.gpu


def _norm_err(X32, X64, rp, col):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    diag = X64[rp[:-1] + np.array([np.searchsorted(col[rp[i]:rp[i + 1]], i) for i in range(n)])]
    scale = np.sqrt(np.abs(diag[rows] * diag[col]))
    scale = np.where(scale > 0, scale, np.abs(X64).max())
    return float((np.abs(X32 - X64) / scale).max())


@pytest.mark.parametrize("case", ["halfspace_block", "hex8_gpml_aniso_topo", "hex20_gpml", "hex27_gpml_tensor_topo"])
def test_fp32_integration_operators(case, need_gpu):
    if case == "halfspace_block":
        m = M.halfspace_block()
    elif case == "hex8_gpml_aniso_topo":
        m = M.random_small(8, 5, 3, 3, seed=33, dirichlet=False, nextd=1, anisotropic=True, topo=200.0, freqs=(1.0,))
    elif case == "hex20_gpml":
        m = M.random_small(20, 4, 4, 4, seed=31, dirichlet=False, nextd=1, anisotropic=True, freqs=(1.0,))
    else:
        m = M.random_small(27, 3, 3, 3, seed=32, dirichlet=False, nextd=1, full_tensor=True, topo=150.0,
                           mu_contrast=True, freqs=(1.0,))
    print(f"fp32 integration {case}", flush=True)
    g = GpuVfem(m)
    rp, col, K64, M64 = g.operators()
    t32 = g.assemble("fp32")
    _, _, K32, M32 = g.operators()
    g.set_frequency(0)
    with pytest.raises(MvfError):
        g.global_vfem()  # the reference's FP64 matrix is not what the handle holds now
    t64 = g.assemble("fp64")
    _, _, K64b, M64b = g.operators()
    assert np.array_equal(K64, K64b) and np.array_equal(M64, M64b)
    ek, em = _norm_err(K32, K64, rp, col), _norm_err(M32, M64, rp, col)
    fk = np.linalg.norm(K32 - K64) / np.linalg.norm(K64)
    fm = np.linalg.norm(M32 - M64) / np.linalg.norm(M64)
    print(f"{case}: K max normalized err {ek:.2e} (Frobenius {fk:.2e}), M {em:.2e} ({fm:.2e}); "
          f"element kernel {t64:.3f} ms fp64, {t32:.3f} ms fp32", flush=True)
    assert ek <= 2e-4 and em <= 2e-4
    assert fk <= 1e-6 and fm <= 1e-6


def test_fp32_integration_solution_error(need_gpu):
    """Config 1: the FP32-integrated operator's solution against the FP64 reference solve."""
    m = M.halfspace_block()
    g = GpuVfem(m)
    g.set_frequency(0)
    ia, ja, a, b = g.global_vfem()
    ia, ja, a = g.remove_zeros(ia, ja, a)
    n = g.nne
    U = sp.coo_matrix((a, (ia - 1, ja - 1)), shape=(n, n)).tocsr()
    A = (U + sp.triu(U, 1).T).tocsc()
    lu = spla.splu(A)
    b2 = b.reshape(2, n)
    xref = np.stack([lu.solve(b2[d]) for d in range(2)])
    mask = determined_dofs(m, g.gne())
    xs, ys = config1_receivers(m)
    ke_top = m.nz - 4
    o = Oracle(m)
    o.set_frequency(0)
    oes, ohs = o.node_solution(xref.ravel())
    _, _, _, rho_ref, phi_ref = o.receivers(oes, ohs, ke_top, xs, ys)
    results = {}
    for prec_int in ("fp64", "fp32"):
        g.assemble(prec_int)
        x, info = g.solve(b, method="gmres", precond="direct", matrix="c64ref", tol=1e-8)
        assert info["converged"], info
        assert info["operator_fp32"] == (prec_int == "fp32")
        x = x.reshape(2, n)
        err = max(np.linalg.norm(x[d][mask] - xref[d][mask]) / np.linalg.norm(xref[d][mask]) for d in range(2))
        out = g.readout(x.ravel(), ke_top=ke_top, rx=xs, ry=ys)
        drho = max(float((np.abs(out[5][:, c] - rho_ref[:, c]) / rho_ref[:, c]).max()) for c in (1, 2))
        dphi = max(float(np.abs(out[6][:, c] - phi_ref[:, c]).max()) for c in (1, 2))
        results[prec_int] = (err, drho, dphi)
        print(f"{prec_int} integration: secondary E rel L2 {err:.2e}, rho_a {drho:.2e}, phase {dphi:.2e} deg, "
              f"{info['iterations']} GMRES iterations", flush=True)
    g.assemble("fp64")
    e64, r64, p64 = results["fp64"]
    e32, r32, p32 = results["fp32"]
    assert e64 <= 1e-6 and r64 <= 1e-3 and p64 <= 0.05
    assert e32 <= 5e-4 and r32 <= 1e-3 and p32 <= 0.05


def test_complex64_factor(need_gpu):
    """The emulated complex64 factorization (MVF_PREC_MIXED with the direct
    preconditioner: every front entry rounded to complex64 after each update)
    as the preconditioner of FP64 GMRES, against the FP64 factor, on a 20-node
    GPML model at the lowest and a middle frequency.  The FP64 factor converges
    in a few iterations.  The complex64 one is recorded: on the 20-node
    64 x 64 x 48 workload it failed at all 8 frequencies (residual 0.53 after
    120 iterations at 1e-3 Hz, NaN from the vanishing air-regularization
    pivots at the others), so no bar is set on it beyond running to a defined
    end (converged, or MVF_E_NOCONV)."""
    m = M.anisotropic_blocks(nx=12, ny=12, nz_earth=8, nz_air=4, nodes=20, freqs=(1e-3, 1.0), block=4, seed=6,
                             nextd=3)
    g = GpuVfem(m)
    for f in m.freqs:
        i64 = g.solve_frequencies([f], method="gmres", precond="direct", tol=1e-8, maxit=60)
        assert i64["converged"] and max(i64["true_resid"][:2]) <= 1e-8 and i64["iterations"] <= 10
        try:
            i32 = g.solve_frequencies([f], method="gmres", precond="direct", precision="mixed", tol=1e-8, maxit=60)
            note = f"converged in {i32['iterations']} iterations"
            assert max(i32["true_resid"][:2]) <= 1e-8
        except MvfError as e:
            assert e.code == L.MVF_E_NOCONV, str(e)[:200]
            note = "did not converge in 60 iterations"
        print(f"f={f:g} Hz: FP64 factor {i64['iterations']} iterations; complex64 factor {note}", flush=True)
    g.close()
