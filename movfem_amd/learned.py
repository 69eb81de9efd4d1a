"""Learned initial guess for the frequency-model solve (host side).

No counterpart in the reference, which solves every frequency from scratch
with MUMPS (MoVFEM_3DMT.f90:113-115).  A 3-D U-Net (libmovfem_gpu.so,
mvf_nn_*: BF16 matrix-core training and inference) maps per-cell features of
an 8-node model on its voxel grid to the model's edge unknowns; the
prediction, scaled per polarization by the least-squares factor that
minimises ||b - alpha A x0|| (mvf_solve_guess), starts the Krylov solve.  The
FP64 true residual judges every result and plain zero-start Krylov remains the
fallback: the scaled guess never starts above ||b||.

Workload (the learned-preconditioner configuration): 8-node Dirichlet models
on 64 x 64 x 48 cells of 500 m (40 earth + 8 air layers), layered earths of
3-6 layers with log10 sigma ~ U[-3, 0] plus 1-4 conductive or resistive
boxes, frequency log-uniform in [1e-3, 1e2] Hz; the training fields come from
the device solver itself.

Voxel layout of the unknowns: an edge along axis d starting at lattice node
(i, j, k) is voxel (i, j, k) of channel group d (i = cell index along d, the
lattice index along the other two).  Every interior edge of a Dirichlet model
fits the cell grid, since lattice index 0 and n along the transverse axes are
boundary edges the reference removes (global_assembly.f90:266-285).
Channels: 6 * pol + 2 * d + (0 real, 1 imaginary).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from . import model as M
from .vfem import GpuVfem, MvfError, _dp, info_dict, solver_opts

# (basis node, direction) of the 8-node element's 12 edges (v_fem.f90:491-504)
_ED12 = [(4, 3), (4, 2), (8, 2), (3, 3), (4, 1), (8, 1), (3, 1), (7, 1), (1, 3), (1, 2), (5, 2), (2, 3)]
# reference coordinates of the 8 corner nodes (n_fem.f90:113-121)
_NREF = [(1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, -1), (1, -1, 1), (1, 1, 1), (-1, 1, 1), (-1, -1, 1)]

CIN, COUT = 3, 12


def training_model(seed: int, nx=64, ny=64, nz_earth=40, nz_air=8, h=500.0) -> tuple:
    """(MTModel, frequency) of one seeded layered earth with boxes (8-node, Dirichlet)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nz = nz_earth + nz_air
    f = float(10.0 ** rng.uniform(-3.0, 2.0))
    xp = np.arange(nx + 1, dtype=np.float64) * h
    yp = np.arange(ny + 1, dtype=np.float64) * h
    zl = np.arange(nz + 1, dtype=np.float64) * h
    m = M.MTModel(nx, ny, nz, 8, xp, yp, np.zeros(1), np.zeros((1, 6), complex), np.zeros((1, 6)),
                  np.asarray([f]), dirichlet=True, name=f"layered{seed}")
    i, j, k = M.node_grid(m)
    m.zp = zl[k].copy()
    m.ztop = nz_earth * h
    air = m.zp > m.ztop + 1e-9
    # layers: interface depths (in cells below the surface) and conductivities
    nl = int(rng.integers(3, 7))
    cuts = np.sort(rng.choice(np.arange(1, nz_earth), size=nl - 1, replace=False))
    lsig = 10.0 ** rng.uniform(-3.0, 0.0, size=nl)
    depth_cells = nz_earth - k  # lattice layers below the surface (node k)
    layer = np.searchsorted(cuts, depth_cells, side="right")
    sig = lsig[np.minimum(layer, nl - 1)]
    for _ in range(int(rng.integers(1, 5))):
        sx, sy, sz = (int(rng.integers(2, 13)), int(rng.integers(2, 13)), int(rng.integers(2, 11)))
        x0 = int(rng.integers(2, nx - sx - 1))
        y0 = int(rng.integers(2, ny - sy - 1))
        z1 = nz_earth - int(rng.integers(1, nz_earth - sz))
        inb = (i >= x0) & (i <= x0 + sx) & (j >= y0) & (j <= y0 + sy) & (k >= z1 - sz) & (k <= z1)
        sig = np.where(inb, 10.0 ** rng.uniform(-3.0, 0.0), sig)
    m.sigma = M.fill_isotropic(sig, air, m.omega(0))
    m.mu = M.mu_free(m.npt)
    m.nextd = 2
    m.nzl_top = 2
    return m, f


def features(m: M.MTModel, f: float) -> np.ndarray:
    """(nx, ny, nz, 3) float32: normalised log10 of the cell conductivity (mean of
    its corner nodes' Re sigma; air cells 0), the air mask and log10 f."""
    s = m.sigma[:, 0].real.reshape(m.nnx, m.nny, m.nnz)
    cell = 0.125 * sum(s[a:a + m.nx, b:b + m.ny, c:c + m.nz] for a in (0, 1) for b in (0, 1) for c in (0, 1))
    air = cell <= 0
    out = np.zeros((m.nx, m.ny, m.nz, CIN), np.float32)
    out[..., 0] = np.where(air, 0.0, (np.log10(np.maximum(cell, 1e-12)) + 1.5) / 1.5)
    out[..., 1] = air
    out[..., 2] = math.log10(f) / 2.5
    return out


def dof_voxels(m: M.MTModel, gne: np.ndarray):
    """For every global DOF (0-based): (direction, i, j, k) of its voxel."""
    ne = m.ne
    ie = np.repeat(np.arange(m.nx), m.ny * m.nz)
    je = np.tile(np.repeat(np.arange(m.ny), m.nz), m.nx)
    ke = np.tile(np.arange(m.nz), m.nx * m.ny)
    nne = int(gne.max())
    d_of = np.full(nne, -1, np.int64)
    ijk = np.zeros((nne, 3), np.int64)
    for im, (node, dr) in enumerate(_ED12):
        g = gne[:, im]
        ok = g > 0
        off = [(c + 1) // 2 for c in _NREF[node - 1]]
        pos = np.stack([ie + off[0], je + off[1], ke + off[2]], axis=1)
        pos[:, dr - 1] -= off[dr - 1]  # cell index along the edge's axis
        d_of[g[ok] - 1] = dr - 1
        ijk[g[ok] - 1] = pos[ok]
    assert (d_of >= 0).all()
    assert (ijk[:, 0] < m.nx).all() and (ijk[:, 1] < m.ny).all() and (ijk[:, 2] < m.nz).all()
    _ = ne
    return d_of, ijk


def determined(m: M.MTModel, gne: np.ndarray) -> np.ndarray:
    """DOFs the system determines: edges of elements with a conducting node.  The
    reference keeps only Re(sigma) in the mass term (integration.f90:211-238)
    and air nodes carry a purely imaginary sigma, so elements whose nodes are all
    air add curl-curl only and gradients supported there are a null space: any
    solver returns an arbitrary gauge on the other DOFs, which the network
    neither learns nor predicts."""
    o = m.nord - 1
    cond = (m.sigma[:, [0, 3, 5]].real > 0).any(axis=1).reshape(m.nnx, m.nny, m.nnz)
    ce = np.zeros((m.nx, m.ny, m.nz), bool)
    for a in range(o + 1):
        for b in range(o + 1):
            for c in range(o + 1):
                ce |= cond[a:a + m.nx * o:o, b:b + m.ny * o:o, c:c + m.nz * o:o]
    g = gne[ce.ravel()]
    out = np.zeros(int(gne.max()), bool)
    out[g[g > 0] - 1] = True
    return out


def field_scale(x: np.ndarray, det: np.ndarray) -> np.ndarray:
    """Per polarization, the RMS of the solution over the determined DOFs: the
    network learns the field's shape, the least-squares factor of
    mvf_solve_guess restores its amplitude and phase."""
    return np.sqrt((np.abs(x[:, det]) ** 2).mean(axis=1))


def to_voxels(m, d_of, ijk, x: np.ndarray, det: np.ndarray):
    """x (2, nne) complex -> target (nx, ny, nz, 12) float32, each polarization
    divided by its RMS over the determined DOFs, and the mask of those DOFs."""
    t = np.zeros((m.nx, m.ny, m.nz, COUT), np.float32)
    mask = np.zeros_like(t)
    s = field_scale(x, det)
    i, j, k = ijk[det, 0], ijk[det, 1], ijk[det, 2]
    for p in range(2):
        for part in range(2):
            ch = 6 * p + 2 * d_of[det] + part
            v = (x[p, det].real if part == 0 else x[p, det].imag) / s[p]
            t[i, j, k, ch] = v
            mask[i, j, k, ch] = 1.0
    return t, mask


def from_voxels(y: np.ndarray, d_of, ijk, det: np.ndarray) -> np.ndarray:
    """prediction (nx, ny, nz, 12) -> x0 (2, nne) complex, zero on the DOFs the
    system leaves undetermined."""
    x = np.zeros((2, len(d_of)), np.complex128)
    for p in range(2):
        re = y[ijk[:, 0], ijk[:, 1], ijk[:, 2], 6 * p + 2 * d_of]
        im = y[ijk[:, 0], ijk[:, 1], ijk[:, 2], 6 * p + 2 * d_of + 1]
        x[p] = np.where(det, re + 1j * im, 0.0)
    return x


def operator(g: GpuVfem, omega: float):
    """The complex64-rounded operator the reference hands MUMPS (K + i fl32(omega) M
    with both parts rounded to float32), as a SciPy CSR matrix."""
    import scipy.sparse as sp

    rp, col, K, Mv = g.operators()
    w32 = float(np.float32(omega))
    a = K.astype(np.float32).astype(np.float64) + 1j * (w32 * Mv).astype(np.float32).astype(np.float64)
    return sp.csr_matrix((a, col, rp), shape=(g.nne, g.nne))


def fit_parts(A, b: np.ndarray, x0: np.ndarray, d_of: np.ndarray) -> tuple:
    """Per polarization, the combination sum_k c_k v_k of six parts of the
    prediction (its real and imaginary parts on the edges of each direction)
    with complex c_k minimising ||b - A sum_k c_k v_k||; returns the guess
    (2, nne) and the initial relative residual of the plain and the fitted guess."""
    out = np.zeros_like(x0)
    r0 = []
    for p in range(2):
        parts = []
        for d in range(3):
            sel = d_of == d
            parts.append(np.where(sel, x0[p].real, 0.0).astype(np.complex128))
            parts.append(np.where(sel, 1j * x0[p].imag, 0.0))
        V = np.stack(parts, axis=1)
        AV = np.column_stack([A @ V[:, k] for k in range(V.shape[1])])
        c, *_ = np.linalg.lstsq(AV, b[p], rcond=None)
        out[p] = V @ c
        a1 = (A @ x0[p]).conj() @ b[p] / max(np.linalg.norm(A @ x0[p]) ** 2, 1e-300)
        r0.append((np.linalg.norm(b[p] - a1 * (A @ x0[p])) / np.linalg.norm(b[p]),
                   np.linalg.norm(b[p] - AV @ c) / np.linalg.norm(b[p])))
    return out, r0


def solve_sample(m: M.MTModel, f: float, g: GpuVfem | None = None):
    """The field the network learns: the device solve of the model (direct-
    preconditioned GMRES to an FP64 true residual <= 1e-8)."""
    g = g or GpuVfem(m)
    info = g.solve_frequencies([f], method="gmres", precond="direct", tol=1e-8)
    if not info["converged"]:
        raise MvfError(L.MVF_E_NOCONV, f"training solve did not converge {info}")
    return g, g.solution(2), info


class Net:
    """The U-Net of libmovfem_gpu.so (mvf_nn_*)."""

    def __init__(self, nx=64, ny=64, nz=48, batch=2, seed=4, lr=1e-3, device=0):
        self._lib = L.lib()
        c = L.NnConfig()
        self._lib.mvf_nn_default_config(C.byref(c))
        c.nx, c.ny, c.nz, c.cin, c.cout, c.batch, c.seed, c.lr, c.device = nx, ny, nz, CIN, COUT, batch, seed, lr, device
        self.cfg = c
        self.h = C.c_void_p()
        rc = self._lib.mvf_nn_create(C.byref(c), C.byref(self.h))
        if rc:
            msg = self._lib.mvf_nn_last_error(self.h).decode() if self.h else ""
            self.close()
            raise MvfError(rc, msg)
        self.batch = batch
        self.shape = (nx, ny, nz)

    def _check(self, rc):
        if rc:
            raise MvfError(rc, self._lib.mvf_nn_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self._lib.mvf_nn_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def num_params(self) -> int:
        return int(self._lib.mvf_nn_num_params(self.h))

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params(), np.float32)
        self._check(self._lib.mvf_nn_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def load(self, path) -> dict:
        """Weights saved by train_learned (BF16 bits) into this network; returns
        the file's other fields (loss history, evaluation)."""
        with np.load(path, allow_pickle=False) as z:
            bits = z["params_bf16"].astype(np.uint32)
            meta = {k: z[k] for k in z.files if k != "params_bf16"}
        self.set_params((bits << 16).view(np.float32))
        return meta

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params()
        self._check(self._lib.mvf_nn_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def grads(self) -> np.ndarray:
        out = np.empty(self.num_params(), np.float32)
        self._check(self._lib.mvf_nn_get_grads(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def forward_backward(self, x, t, mask=None) -> float:
        fp = C.POINTER(C.c_float)
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        mp = None if mask is None else np.ascontiguousarray(mask, np.float32)
        loss = C.c_float()
        self._check(self._lib.mvf_nn_forward_backward(self.h, self.batch, x.ctypes.data_as(fp), t.ctypes.data_as(fp),
                                                      None if mp is None else mp.ctypes.data_as(fp), C.byref(loss)))
        return float(loss.value)

    def adam(self, lr: float):
        self._check(self._lib.mvf_nn_adam(self.h, float(lr)))

    def predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty(x.shape[:-1] + (COUT,), np.float32)
        fp = C.POINTER(C.c_float)
        self._check(self._lib.mvf_nn_predict(self.h, self.batch, x.ctypes.data_as(fp), y.ctypes.data_as(fp)))
        return y

    def set_timing(self, on: bool):
        self._check(self._lib.mvf_nn_set_timing(self.h, int(on)))

    def stats(self) -> dict:
        s = np.zeros(3)
        self._check(self._lib.mvf_nn_stats(self.h, _dp(s)))
        return {"conv_ms": s[0], "conv_flops": s[1], "conv_launches": int(s[2])}


def solve_with_guess(g: GpuVfem, b: np.ndarray, x0: np.ndarray, **kw):
    """mvf_solve_guess: the zmumps-shaped solve started from x0 (2 * nne complex)."""
    x = np.ascontiguousarray(b, np.complex128).copy()
    x0 = np.ascontiguousarray(x0, np.complex128).ravel()
    o = solver_opts(**kw)
    info = L.SolveInfo()
    alpha = np.zeros(4)
    rc = g._lib.mvf_solve_guess(g.h, _dp(x0.view(np.float64)), _dp(x.view(np.float64)), C.byref(o), C.byref(info),
                                _dp(alpha))
    d = info_dict(info, 2)
    d["alpha"] = [complex(alpha[0], alpha[1]), complex(alpha[2], alpha[3])]
    if rc:
        raise MvfError(rc, g._lib.mvf_last_error(g.h).decode() + f" {d}")
    return x, d
