"""Benchmark of the element-integration -> assembly -> solve path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload aniso20|halfspace8]

Workload (default, aniso20): 20-node edge elements on a 64 x 64 x 48 cell grid
(40 earth + 8 air layers of 500 m), 8 frequencies log-spaced 1e-3 .. 100 Hz,
diagonal anisotropic sigma (log10 of each axis uniform in [-3, 0] per 4x4x4
cell block, PCG64 seed 2), GPML (Fang scheme, a0 = b0 = 1, N = 2) on the +x, +y
and top 4 cells.  Inputs are synthetic and resident in HBM before timing.

One step = re-integrate every element's K and M on the device and re-assemble
them (the reference's per-frequency element loop), then for each of the 8
frequencies: device RHS, numeric multifrontal LU factorization (FP64, nested
dissection) of the operator and right-preconditioned FP64 GMRES to a true
relative residual <= 1e-8 per polarization (the reference factors every
frequency with MUMPS, MoVFEM_3DMT.f90:113-115).  The value is
frequency-model solves per second (one solve = one frequency, both
polarizations).  The plain GPU Krylov path (Jacobi COCG, FP64) is attempted
at the lowest and highest frequency for comparison; on this GPML model it
does not reach 1e-8.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "freq-model solves/s @ rel resid 1e-8 (learned+mixed vs plain GPU); elem mats/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
FP32_PEAK_TFS = 157.3
# algorithmic flops per element matrix (K, M upper and b), canonical counts
ELEM_FLOPS = {8: 21768, 20: 376731, 27: 744795}
# HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE runs, each scaled
# by what the counter reads for a plain copy of known size (profiles/)
TRAFFIC_MEASURED = {("halfspace8", "k_spmv"): 20.17e6, ("aniso20", "k_fa_gemm"): 3.43e9}


def workload(name):
    from movfem_amd import model as M

    if name == "aniso20":
        m = M.anisotropic_blocks(nx=64, ny=64, nz_earth=40, nz_air=8, h=500.0, nodes=20,
                                 freqs=tuple(np.logspace(-3, 2, 8)), block=4, seed=2, nextd=4)
        desc = ("20-node edge elements, 64x64x48 cells (40 earth + 8 air, 500 m), 8 freqs logspace(1e-3,100) Hz, "
                "diagonal anisotropic sigma log-uniform [1e-3,1] S/m per 4^3 block (seed 2), GPML Fang a0=b0=1 N=2")
    elif name == "halfspace8":
        m = M.halfspace_block()
        desc = ("8-node edge elements, 20x20x20 cells (16 earth + 4 air, 500 m), 0.01 S/m half-space with a 1 S/m "
                "2x2x1 km block, 1 Hz, Dirichlet")
    else:
        raise SystemExit(f"unknown workload {name}")
    return m, desc


def cpu_baseline(m_full, budget_s=25.0):
    """Reference path on the host: the reference's own modules (oracle/_ref) for
    element integration + assembly + sort of one frequency, SciPy SuperLU
    (complex128, stand-in for MUMPS 4.10) for the 2-RHS solve.  A workload
    small enough to finish in the budget is timed whole; a larger one on a
    reduced model of the same family, scaled linearly to the full model."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from movfem_amd import model as M
    from oracle.oracle import Oracle, ref_available

    nodes = m_full.nodes_per_elem
    if m_full.ne <= 10000:
        m = m_full
    else:
        n = 10 if nodes == 20 else 16
        m = M.anisotropic_blocks(nx=n, ny=n, nz_earth=max(2, (3 * n) // 4 - 2), nz_air=2, h=500.0, nodes=nodes,
                                 freqs=(1e-3,), block=4, seed=2, nextd=2)
    kind = "reference" if ref_available() else "port"
    t0 = time.perf_counter()
    if kind == "reference":
        from oracle.refrun import run_reference

        t_setup = time.perf_counter()
        r = run_reference(m, vfem_freqs=[1])
        ia, ja, a, b = r["ia_1"], r["ja_1"], r["a_1"], r["b_1"]
        nne = int(r["nne"])
        t_asm = time.perf_counter() - t_setup
    else:
        o = Oracle(m)
        o.set_frequency(0)
        t_setup = time.perf_counter()
        ia, ja, a, b = o.global_vfem()
        nne = o.nne
        t_asm = time.perf_counter() - t_setup
    keep = a != 0
    U = sp.coo_matrix((a[keep], (ia[keep] - 1, ja[keep] - 1)), shape=(nne, nne)).tocsr()
    A = (U + sp.triu(U, 1).T).tocsc()
    t1 = time.perf_counter()
    lu = spla.splu(A)
    for d in range(2):
        lu.solve(b[d * nne:(d + 1) * nne])
    t_solve = time.perf_counter() - t1
    what = f"{'reference modules' if kind == 'reference' else 'C restatement'}"
    if m is m_full:
        t_full = t_asm + t_solve
        sample = (f"the whole workload ({m.ne} elements, {nne} DOFs), one frequency: {what} element loop + assembly + "
                  f"sort {t_asm:.2f}s, SciPy SuperLU (MUMPS stand-in) factor + 2 solves {t_solve:.2f}s")
    else:
        full_ne, full_nne = m_full.ne, m_full.nne_hint
        t_full = t_asm * (full_ne / m.ne) + t_solve * (full_nne / nne)
        sample = (f"{m.nx}x{m.ny}x{m.nz} {nodes}-node model of the same family ({m.ne} elements, {nne} DOFs): {what} "
                  f"element loop + assembly + sort {t_asm:.2f}s, SciPy SuperLU (MUMPS stand-in) factor + 2 solves "
                  f"{t_solve:.2f}s; scaled linearly to {full_ne} elements / {full_nne} DOFs (linear scaling flatters "
                  f"the direct solve)")
    return {"value": 1.0 / t_full, "unit": "solves/s", "cores": 1, "kind": kind, "sample": sample,
            "wall_s": time.perf_counter() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="aniso20")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--maxit", type=int, default=3000)
    ap.add_argument("--nfreq", type=int, default=0, help="solve only the first N frequencies (profiling runs)")
    ap.add_argument("--direct-tile", type=int, default=0, help="factorization GEMM tile: 0 by front size, 64 or 128")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB three times (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mib:
        import torch

        n = args.copy_probe_mib << 20
        x = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    from movfem_amd import _lib as L
    from movfem_amd.vfem import GpuVfem, MvfError

    m, desc = workload(args.workload)
    freqs = list(m.freqs)[:args.nfreq] if args.nfreq > 0 else list(m.freqs)
    # The reference's update_sigma refreshes only the first npt+30 linear
    # entries of sigma (geometry.f90:146-152): from the second frequency on,
    # air keeps the first frequency's i*omega*eps and the right-hand side gets a
    # component the air null space cannot absorb, which caps the attainable
    # relative residual above 1e-8.  The timed solves use the corrected
    # update; every other reference semantics stays emulated.
    g = GpuVfem(m, emulate=L.MVF_EMU_GPML_PREV_FLAGS)
    print(f"# setup nne={g.nne} nnz={g.nnze} times={g.setup_times()}", file=sys.stderr, flush=True)

    def step():
        ms = g.assemble("fp64")
        info = g.solve_frequencies(freqs, method="gmres", precond="direct", tol=1e-8, direct_tile=args.direct_tile)
        return ms, info

    for w in range(args.warmup):
        t_w = time.perf_counter()
        step()
        print(f"# warmup step {w} {time.perf_counter() - t_w:.2f}s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    infos, elem_ms = [], []
    for k_ in range(args.steps):
        ms, info = step()
        elem_ms.append(ms)
        infos.append(info)
        print(f"# step {k_} its={info['iterations']} factor={info['factor_seconds']:.2f}s "
              f"resid={max(info['true_resid'][:2 * len(freqs)]):.2e}", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    value = len(freqs) * args.steps / dt
    last = infos[-1]
    # dominant kernel: the factorization's batched complex GEMM
    gemm_tfs = last["gemm_flops"] / (last["gemm_ms"] * 1e-3) / 1e12 if last["gemm_ms"] > 0 else 0.0
    roofline = {"bound": "mfma", "kernel": "k_fa_gemm_mfma (FP64 complex on v_mfma_f64_16x16x4, three real products per complex one on the "
                "128x64 tiles; achieved counts 8 real flops per complex multiply-add; MI355X FP64 matrix peak = FP64 "
                "vector peak; launch durations include time shared with the panel kernels the lookahead overlaps)", "achieved": gemm_tfs, "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                "frac": gemm_tfs / FP64_PEAK_TFS, "traffic": TRAFFIC_MEASURED.get((args.workload, "k_fa_gemm")),
                "avg_launch_ms": last["gemm_ms"] / max(1, last["gemm_launches"]),
                "launches": last["gemm_launches"], "flops_per_factorization": last["factor_flops"],
                "factor_seconds_per_frequency": last["factor_seconds"] / len(freqs),
                "factor_tflops_all_kernels": last["factor_flops"] * len(freqs) / max(last["factor_seconds"], 1e-9)
                / 1e12}
    k = int(np.argmax(last["spmv_ms"]))
    avg_ms = last["spmv_ms"][k] / max(1, last["spmv_launches"][k])
    spmv = {"bound": "hbm", "kernel": "k_spmv<double>", "achieved": last["spmv_bytes"][k] / (avg_ms * 1e-3) / 1e9
            if avg_ms > 0 else 0.0, "peak": HBM_PEAK_GBS, "unit": "GB/s", "avg_launch_ms": avg_ms,
            "launches": last["spmv_launches"][k], "bytes_per_launch": last["spmv_bytes"][k],
            "traffic": TRAFFIC_MEASURED.get((args.workload, "k_spmv"))}
    spmv["frac"] = spmv["achieved"] / HBM_PEAK_GBS
    # element kernel (FP64 and FP32 integration)
    e64 = float(np.median(elem_ms))
    e32 = g.assemble("fp32")
    g.assemble("fp64")
    F = ELEM_FLOPS[m.nodes_per_elem]
    elem = {"elem_mats_per_s": m.ne / (e64 * 1e-3), "elem_mats_per_s_fp32": m.ne / (e32 * 1e-3),
            "elem_kernel_ms": e64, "elem_kernel_ms_fp32": e32,
            "elem_roofline": {"bound": "fp64 valu", "achieved": m.ne * F / (e64 * 1e-3) / 1e12, "peak": FP64_PEAK_TFS,
                              "unit": "TFLOP/s", "frac": m.ne * F / (e64 * 1e-3) / 1e12 / FP64_PEAK_TFS,
                              "flops_per_elem": F}}
    out = {"metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "c128 (FP64 complex factorization and GMRES)",
           "data": "synthetic", "config": {"workload": args.workload, "description": desc, "elements": m.ne,
                                           "nne": g.nne, "nnz": g.nnze, "frequencies": len(freqs),
                                           "path": "device multifrontal LU (FP64, nested dissection) preconditioning "
                                                   "FP64 GMRES, one factorization per frequency"},
           "iterations": last["iterations"], "precond_applies": last["precond_applies"],
           "true_resid_max": max(last["true_resid"][:2 * len(freqs)]), "roofline": roofline, "spmv_roofline": spmv}
    out.update(elem)
    if not args.no_plain:
        # plain GPU Krylov (Jacobi COCG, FP64) at the lowest and highest frequency
        plain = []
        for f in (freqs[0], freqs[-1]):
            t1 = time.perf_counter()
            try:
                inf = g.solve_frequencies([f], precision="fp64", tol=1e-8, maxit=args.maxit, check_every=50)
                ok, res, its = True, max(inf["true_resid"][:2]), inf["iterations"]
            except MvfError as e:
                ok, res, its = False, None, args.maxit
                print(f"# plain COCG at {f} Hz: {str(e)[:160]}", file=sys.stderr, flush=True)
            plain.append({"freq_hz": f, "converged": ok, "iterations": its, "true_resid": res,
                          "seconds": time.perf_counter() - t1})
        out["plain_gpu"] = plain
        if all(p["converged"] for p in plain):
            out["plain_gpu_solves_per_s"] = len(plain) / sum(p["seconds"] for p in plain)
            out["speedup_vs_plain"] = value / out["plain_gpu_solves_per_s"]
        else:
            out["plain_gpu_solves_per_s"] = 0.0
            out["speedup_vs_plain"] = None
    if not args.no_cpu:
        m.nne_hint = g.nne
        cb = cpu_baseline(m)
        out["cpu_baseline"] = {k2: cb[k2] for k2 in ("value", "unit", "cores", "kind", "sample")}
        out["speedup_vs_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
