"""Train the learned initial guess on solver-generated fields and measure its gain.

    python -m movfem_amd.train_learned --samples N --epochs E --out weights.npz [--heldout K]

1. N seeded models of the learned-preconditioner workload (movfem_amd.learned:
   8-node, 64 x 64 x 48 cells, layered earth with boxes, Dirichlet, one
   frequency each, log-uniform in [1e-3, 1e2] Hz), each solved on the device
   (direct-preconditioned GMRES, FP64 true residual <= 1e-8): inputs are the
   per-cell features, targets the solution's edge unknowns on the voxel grid.
2. Adam on the U-Net (BF16 matrix cores, FP32 master weights), batch 2,
   linear warm-up over the first --warmup steps, then cosine decay.
3. K held-out models (seeds not used in training): Jacobi COCG to 1e-8 from
   zero and from the network's scaled prediction; iterations and time of both.
The weights (BF16 bits), the loss history and the evaluation go to --out; a
JSON summary line goes to stdout.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import learned as Lm


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def bf16_bits(p: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(p, np.float32).view(np.uint32)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)


def make_sample(seed: int):
    m, f = Lm.training_model(seed)
    g, x, info = Lm.solve_sample(m, f)
    gne = g.gne()
    d_of, ijk = Lm.dof_voxels(m, gne)
    t, mask = Lm.to_voxels(m, d_of, ijk, x, Lm.determined(m, gne))
    feat = Lm.features(m, f)
    g.close()
    return feat, t, mask, info


def evaluate(net: Lm.Net, seeds, maxit=40000, tol=1e-8):
    """Jacobi COCG from zero and from the scaled prediction on held-out models."""
    rows = []
    for s in seeds:
        m, f = Lm.training_model(s)
        g = Lm.GpuVfem(m)
        g.set_frequency_hz(f)
        _, _, _, b = g.global_vfem()
        gne = g.gne()
        d_of, ijk = Lm.dof_voxels(m, gne)
        feat = Lm.features(m, f)
        batch = np.stack([feat] * net.batch)
        t0 = time.perf_counter()
        y = net.predict(batch)[0]
        t_pred = time.perf_counter() - t0
        x0 = Lm.from_voxels(y, d_of, ijk, Lm.determined(m, gne))
        A = Lm.operator(g, 2.0 * math.pi * f)
        t1 = time.perf_counter()
        x6, r0 = Lm.fit_parts(A, b.reshape(2, -1), x0, d_of)
        t_fit = time.perf_counter() - t1
        t1 = time.perf_counter()
        _, i0 = g.solve(b, method="cocg", tol=tol, maxit=maxit, check_every=50)
        t2 = time.perf_counter()
        _, i1 = Lm.solve_with_guess(g, b, x0, method="cocg", tol=tol, maxit=maxit, check_every=50)
        t3 = time.perf_counter()
        _, i2 = Lm.solve_with_guess(g, b, x6, method="cocg", tol=tol, maxit=maxit, check_every=50)
        t4 = time.perf_counter()
        rows.append({"seed": s, "freq_hz": f, "nne": g.nne, "its_zero": i0["iterations"], "its_guess": i1["iterations"],
                     "its_fit": i2["iterations"], "s_zero": t2 - t1, "s_guess": t3 - t2 + t_pred,
                     "s_fit": t4 - t3 + t_pred + t_fit, "predict_s": t_pred, "fit_s": t_fit,
                     "r0_guess": [a for a, _ in r0], "r0_fit": [c for _, c in r0],
                     "resid_zero": max(i0["true_resid"][:2]), "resid_guess": max(i1["true_resid"][:2]),
                     "resid_fit": max(i2["true_resid"][:2]), "alpha": [[a.real, a.imag] for a in i1["alpha"]]})
        log(f"held-out {s}: f={f:.3g} its {i0['iterations']} -> {i1['iterations']} (scaled), {i2['iterations']} "
            f"(six-part fit); initial residual {max(a for a, _ in r0):.2e} / {max(c for _, c in r0):.2e}")
        g.close()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--heldout", type=int, default=4)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--budget-s", type=float, default=900.0, help="stop generating / training past this wall time")
    ap.add_argument("--out", default="learned_weights.npz")
    ap.add_argument("--seed", type=int, default=4)
    args = ap.parse_args()
    t_start = time.perf_counter()
    feats, targets, masks, gen_s = [], [], [], []
    for s in range(args.samples):
        if time.perf_counter() - t_start > 0.45 * args.budget_s:
            log(f"generation budget reached after {s} samples")
            break
        t0 = time.perf_counter()
        f_, t_, m_, info = make_sample(s)
        gen_s.append(time.perf_counter() - t0)
        feats.append(f_)
        targets.append(t_)
        masks.append(m_)
        if s % 8 == 0:
            log(f"sample {s}: {gen_s[-1]:.2f}s, target rms {np.sqrt((t_[m_ > 0] ** 2).mean()):.3f}")
    n = len(feats)
    X = np.stack(feats)
    T = np.stack(targets)
    Mk = np.stack(masks)
    net = Lm.Net(batch=2, seed=args.seed, lr=args.lr)
    net.set_timing(True)
    steps_per_epoch = n // 2
    total = steps_per_epoch * args.epochs
    rng = np.random.default_rng(args.seed)
    hist, conv_ms, conv_fl, step_s = [], 0.0, 0.0, []
    k = 0
    t_train = time.perf_counter()
    for ep in range(args.epochs):
        order = rng.permutation(n)
        el = []
        for b0 in range(0, steps_per_epoch * 2, 2):
            if time.perf_counter() - t_start > 0.85 * args.budget_s:
                break
            idx = order[b0:b0 + 2]
            lr = 0.5 * args.lr * (1.0 + math.cos(math.pi * k / max(total, 1))) * min(1.0, (k + 1) / args.warmup)
            ts = time.perf_counter()
            loss = net.forward_backward(X[idx], T[idx], Mk[idx])
            net.adam(lr)
            step_s.append(time.perf_counter() - ts)
            st = net.stats()
            conv_ms += st["conv_ms"]
            conv_fl += st["conv_flops"]
            el.append(loss)
            if k < 400 and k % 25 == 0:
                log(f"step {k}: loss {loss:.4f} lr {lr:.2e}")
            k += 1
        if el:
            hist.append(float(np.mean(el)))
            log(f"epoch {ep}: loss {hist[-1]:.4f} ({len(el)} steps)")
        if time.perf_counter() - t_start > 0.85 * args.budget_s:
            log("training budget reached")
            break
    train_s = time.perf_counter() - t_train
    rows = evaluate(net, [100000 + i for i in range(args.heldout)])
    p = net.params()
    np.savez_compressed(args.out, params_bf16=bf16_bits(p), loss=np.asarray(hist), samples=n, epochs=len(hist),
                        eval=json.dumps(rows), nx=64, ny=64, nz=48, seed=args.seed)
    summ = {"samples": n, "gen_s_mean": float(np.mean(gen_s)), "epochs": len(hist), "steps": k,
            "loss_first": hist[0] if hist else None, "loss_last": hist[-1] if hist else None,
            "train_s": train_s, "step_ms_mean": float(np.mean(step_s) * 1e3) if step_s else None,
            "conv_tflops": conv_fl / max(conv_ms, 1e-9) / 1e9, "eval": rows,
            "mean_its_zero": float(np.mean([r["its_zero"] for r in rows])) if rows else None,
            "mean_its_guess": float(np.mean([r["its_guess"] for r in rows])) if rows else None,
            "mean_its_fit": float(np.mean([r["its_fit"] for r in rows])) if rows else None,
            "wall_s": time.perf_counter() - t_start}
    print(json.dumps(summ), flush=True)


if __name__ == "__main__":
    main()
