"""Drives one kernel family for counter collection (rocprofv3 --pmc / --kernel-trace).

    python -m movfem_amd.profile_kernels elem  [--nodes 20] [--precision fp64|fp32] [--reps 5]
    python -m movfem_amd.profile_kernels spmv  [--nodes 20] [--iters 50]
    python -m movfem_amd.profile_kernels unet  [--steps 3]
    python -m movfem_amd.profile_kernels copy  [--mib 1024]
    python -m movfem_amd.profile_kernels factor [--nodes 20]
    python -m movfem_amd.profile_kernels boundary [--nodes 20] [--reps 2]

elem: mvf_assemble (element integration and assembly of K and M) repeated on
  the anisotropic GPML workload of the given element type (20-node: 64 x 64 x 48
  cells; 8-node: the same grid; 27-node: 32 x 32 x 24 cells; with --full the
  27-node 128 x 128 x 96 topography model, 38M DOFs).
spmv: Jacobi COCG on the same model at 1 Hz for a fixed iteration count (the
  operator apply of every Krylov path, two columns); the solve is stopped by its
  iteration cap, not converged.
unet: forward + backward + Adam steps of the learned-guess network on a batch
  of two 64 x 64 x 48 samples (random inputs).
copy: three device-to-device copies of a known size, the reference the HBM
  traffic counters are scaled by.
factor: one direct-preconditioned GMRES solve at 1e-3 Hz on the elem workload
  with the library's verbose timing on: the factorization's cumulative time
  after each tree level and its GEMM time per mode go to stderr.
boundary: one frequency (1e-3 Hz) through the host-buffer boundary the
  reference's call sequence uses (global_vfem into host ia/ja/a/b, then the
  zmumps(job=6)-shaped mvf_solve on a host right-hand side, solution back to
  the host; then find_zeros/rem_zeros on the host arrays, and the fused
  mvf_global_vfem_nz that drops the zero entries on the device), timed per
  call with the bytes that cross PCIe, beside the same
  frequency solved with the right-hand side and the solution kept on the device
  (mvf_solve_frequencies).  The host arrays are allocated once; the first
  repetition also pays their first-touch page faults and the solver's setup.
Each prints one JSON line with what it ran.
"""
from __future__ import annotations

import argparse
import json

import numpy as np


def model(nodes: int, full: bool = False, freq: float = 1.0):
    from . import model as M

    if full:
        # the 27-node workload at full size: 128 x 128 x 96 cells under Gaussian hills, full-tensor sigma
        return M.hills_tensor()
    if nodes == 27:
        return M.anisotropic_blocks(nx=32, ny=32, nz_earth=20, nz_air=4, h=500.0, nodes=27, freqs=(freq,), block=4,
                                    seed=2, nextd=4)
    return M.anisotropic_blocks(nx=64, ny=64, nz_earth=40, nz_air=8, h=500.0, nodes=nodes, freqs=(freq,), block=4,
                                seed=2, nextd=4)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["elem", "spmv", "unet", "copy", "factor", "boundary"])
    ap.add_argument("--nodes", type=int, default=20)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--full", action="store_true", help="elem: the 27-node 128 x 128 x 96 topography model")
    a = ap.parse_args()
    out = {"what": a.what}
    if a.what == "copy":
        import torch

        n = a.mib << 20
        x = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        out["bytes_per_copy"] = n
    elif a.what == "unet":
        from .learned import Net

        net = Net(batch=2)
        rng = np.random.default_rng(0)
        x = rng.standard_normal((2, 64, 64, 48, 3)).astype(np.float32)
        t = rng.standard_normal((2, 64, 64, 48, 12)).astype(np.float32)
        net.set_timing(True)
        for _ in range(a.steps):
            out["loss"] = net.forward_backward(x, t)
            net.adam(1e-3)
        out.update(net.stats())
        out["tflops"] = out["conv_flops"] / out["conv_ms"] / 1e9
    elif a.what == "boundary":
        import ctypes as C
        import time

        from . import _lib as L
        from .vfem import GpuVfem

        m = model(a.nodes, a.full, freq=1e-3)
        g = GpuVfem(m, emulate=L.MVF_EMU_GPML_PREV_FLAGS)
        out.update(nodes=a.nodes, elements=m.ne, nne=g.nne, nnz=g.nnze, freq_hz=1e-3)
        host, dev = [], []
        bufs = None
        for r in range(a.reps):
            t0 = time.perf_counter()
            g.set_frequency(0)
            # host arrays allocated once and filled again, as a caller's frequency loop does:
            # the first repetition pays their first-touch page faults
            ia, ja, av, b = bufs = g.global_vfem(out=bufs)
            t1 = time.perf_counter()
            x, info = g.solve(b, method="gmres", precond="direct", tol=1e-8)
            t2 = time.perf_counter()
            # the reference's find_zeros/rem_zeros on the host arrays (in place), then the fused
            # device form that drops the zero entries before they cross PCIe
            n = C.c_int64()
            t4 = time.perf_counter()
            L.lib().mvf_remove_zeros(len(ia), ia.ctypes.data_as(C.POINTER(C.c_int32)),
                                     ja.ctypes.data_as(C.POINTER(C.c_int32)),
                                     av.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), C.byref(n))
            t5 = time.perf_counter()
            g.set_frequency(0)
            nz = g.global_vfem_nz()
            t6 = time.perf_counter()
            host.append({"global_vfem_s": t1 - t0, "solve_s": t2 - t1, "total_s": t2 - t0,
                         "remove_zeros_s": t5 - t4, "global_vfem_nz_s": t6 - t5, "nonzeros": len(nz[0]),
                         "iterations": info["iterations"], "true_resid": info["true_resid"][:2]})
            del nz
            t3 = time.perf_counter()
            d = g.solve_frequencies([1e-3], method="gmres", precond="direct", tol=1e-8)
            dev.append({"total_s": time.perf_counter() - t3, "iterations": d["iterations"],
                        "true_resid": d["true_resid"][:2]})
            print(f"boundary rep {r}: host {host[-1]} device {dev[-1]}", flush=True)
        out.update(host_boundary=host, device_resident=dev,
                   bytes_to_host=ia.nbytes + ja.nbytes + av.nbytes + b.nbytes + x.nbytes, bytes_to_device=b.nbytes)
    else:
        from .vfem import GpuVfem, MvfError

        m = model(a.nodes, a.full)
        g = GpuVfem(m, verbose=1 if a.what == "factor" else 0)
        out.update(nodes=a.nodes, elements=m.ne, nne=g.nne, nnz=g.nnze)
        if a.what == "factor":
            info = g.solve_frequencies([1e-3], method="gmres", precond="direct", tol=1e-8)
            out.update(iterations=info["iterations"], true_resid=info["true_resid"][:2],
                       factor_seconds=info.get("factor_seconds"))
        elif a.what == "elem":
            ms = [g.assemble(a.precision) for _ in range(a.reps)]
            out.update(precision=a.precision, reps=a.reps, kernel_ms=ms, elem_per_s=m.ne / (np.median(ms) * 1e-3))
        else:
            try:
                info = g.solve_frequencies([1.0], method="cocg", tol=1e-14, maxit=a.iters, check_every=a.iters)
            except MvfError as e:
                info = {"stopped": str(e)[:120]}
            out.update(iters=a.iters, info={k: info[k] for k in ("spmv_launches", "spmv_ms", "spmv_bytes", "spmv_lpr")
                                              if k in info})
    print(json.dumps(out, default=float), flush=True)


if __name__ == "__main__":
    main()
