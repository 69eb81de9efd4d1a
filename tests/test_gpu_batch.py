"""Batched models x frequencies, graph-captured Krylov iterations and the SpMV
autotuner (no counterpart in the reference, which runs one model through its
frequency loop, MoVFEM_3DMT.f90:62-143).

Reduced batch of the many-model configuration: six models of uneven sizes,
8- and 20-node elements mixed, full-tensor sigma (principal values
log-uniform in [1e-3, 1] S/m rotated by seeded ZYZ Euler angles per 4x4x4
block), Dirichlet boundaries, 4 frequencies log-spaced 1e-3 .. 1e2 Hz.
Bars: every column converged to an FP64 true residual <= 1e-8, and the
batched solutions equal per-model solves of the same models to 1e-9
(relative 2-norm): both run the same kernels on the same data, except that
each handle's SpMV autotuner times its own lane width, and a different width
sums the row products in another order (measured 1.5e-10).
"""
import numpy as np
import pytest

from movfem_amd import _lib as L
from movfem_amd import model as M
from movfem_amd.vfem import GpuVfem, solve_batch

pytestmark = pytest.mark.gpu

FREQS = tuple(np.logspace(-3, 2, 4))
SPECS = [(8, 10, 8, 6), (20, 6, 6, 5), (8, 14, 12, 8), (20, 8, 7, 6), (8, 12, 14, 10), (20, 5, 6, 4)]


def batch_models():
    return [M.anisotropic_blocks(nx=nx, ny=ny, nz_earth=nz, nz_air=2, nodes=nodes, freqs=FREQS, block=4, seed=30 + i,
                                 full_tensor=True, dirichlet=True, nextd=2)
            for i, (nodes, nx, ny, nz) in enumerate(SPECS)]


def test_batch_equals_per_model(need_gpu):
    ms = batch_models()
    print("batch: building", flush=True)
    emu = L.MVF_EMU_GPML_PREV_FLAGS
    gs = [GpuVfem(m, emulate=emu) for m in ms]
    infos, st = solve_batch(gs, FREQS, nthreads=6, method="gmres", precond="direct", tol=1e-8)
    assert (st == 0).all()
    for i, (m, g, inf) in enumerate(zip(ms, gs, infos)):
        assert inf["converged"] and max(inf["true_resid"][:8]) <= 1e-8, (i, inf["true_resid"][:8])
        xb = g.solution(8)
        g1 = GpuVfem(m, emulate=emu)
        info1 = g1.solve_frequencies(list(FREQS), method="gmres", precond="direct", tol=1e-8)
        x1 = g1.solution(8)
        d = np.linalg.norm(xb - x1) / np.linalg.norm(x1)
        print(f"model {i} ({m.nodes_per_elem}-node {m.nx}x{m.ny}x{m.nz}, {g.nne} DOFs): batch vs single {d:.2e}, "
              f"iterations {inf['iterations']}/{info1['iterations']}", flush=True)
        assert d <= 1e-9
        g1.close()


def test_graph_matches_eager(need_gpu):
    """COCG and BiCGStab with the iteration block replayed from a HIP graph give
    the eager solve's iterates: same kernels, same arguments, same order.
    Graph solves alternate with eager ones on the same handle, each graph
    replayed many times between host reads of the convergence state."""
    m = M.random_small(8, 5, 5, 4, seed=41, nz_air=1, freqs=(1.0,))
    g = GpuVfem(m)
    g.set_frequency(0)
    _, _, _, b = g.global_vfem()
    for method in ("cocg", "bicgstab"):
        xe, ie = g.solve(b, method=method, tol=1e-9, check_every=20)
        for rep in range(3):
            xg, ig = g.solve(b, method=method, tol=1e-9, check_every=20, use_graph=True)
            assert ig["graph_launches"] > 0 and ig["iterations"] == ie["iterations"]
            assert np.array_equal(xg, xe), (method, rep, np.abs(xg - xe).max())
            xe2, _ = g.solve(b, method=method, tol=1e-9, check_every=20)
            assert np.array_equal(xe2, xe)
        print(f"{method}: {ie['iterations']} iterations, {ig['graph_launches']} graph replays, "
              f"SpMV lanes per row {ig['spmv_lpr']}", flush=True)
        assert ig["spmv_lpr"] in (4, 8, 16, 32)


def test_direct_sweeps_graph_matches_eager(need_gpu):
    """Direct-preconditioned GMRES with every factor application replayed from
    a HIP graph gives the eager iterates bit for bit: the same kernels on the
    same data, with input and output copied through fixed buffers.  Two
    frequencies, so the second replay follows a refactorization into the same
    buffers; a model whose fronts use the lookahead."""
    m = M.anisotropic_blocks(nx=14, ny=14, nz_earth=9, nz_air=2, nodes=8, freqs=(1e-2, 10.0), block=4, seed=43,
                             full_tensor=True, dirichlet=True, nextd=2)
    g = GpuVfem(m)
    for f in m.freqs:
        ie = g.solve_frequencies([f], method="gmres", precond="direct", tol=1e-10)
        xe = g.solution(2)
        ig = g.solve_frequencies([f], method="gmres", precond="direct", tol=1e-10, use_graph=True)
        xg = g.solution(2)
        print(f"f={f:g} Hz: {ie['iterations']} GMRES iterations eager, {ig['iterations']} with graphs", flush=True)
        assert ig["iterations"] == ie["iterations"] and np.array_equal(xg, xe)
    g.close()


def test_persistent_cocg(need_gpu):
    """COCG with check_every iterations per cooperative launch of the persistent
    kernel (three grid barriers per iteration) against the per-kernel
    iteration: both reach the FP64 true residual bar and agree to 1e-8; the
    block partial sums differ in grouping, so the iterates are not identical."""
    for m in (M.random_small(8, 5, 5, 4, seed=41, nz_air=1, freqs=(1.0,)),
              M.anisotropic_blocks(nx=16, ny=16, nz_earth=10, nz_air=2, nodes=8, freqs=(10.0,), block=4, seed=44,
                                   full_tensor=True, dirichlet=True, nextd=2)):
        g = GpuVfem(m)
        g.set_frequency(0)
        _, _, _, b = g.global_vfem()
        xe, ie = g.solve(b, method="cocg", tol=1e-10, check_every=25, maxit=20000)
        xp, ip = g.solve(b, method="cocg", tol=1e-10, check_every=25, maxit=20000, use_graph="persistent")
        d = np.linalg.norm(xp - xe) / np.linalg.norm(xe)
        print(f"nne={g.nne}: {ie['iterations']} iterations per kernel, {ip['iterations']} persistent "
              f"({ip['graph_launches']} launches); difference {d:.1e}", flush=True)
        assert ip["converged"] and max(ip["true_resid"][:2]) <= 1e-10 and ip["graph_launches"] > 0
        assert d <= 1e-8
        g.close()
