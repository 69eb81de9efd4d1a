"""Host-side mirror of the reference's interfaces for the hot path.

The reference drives the path from its frequency loop (MoVFEM_3DMT.f90:62-143):

    initialize_vfem()                       -> GpuVfem(model)
    update_omega / update_sigma / pset_pmodel -> set_frequency(i)
    int_elem_params + alocal/blocal         -> integrate(e0, e1), alocal(e), blocal(e)
    global_vfem(irn, jcn, a, rhs)           -> global_vfem()
    find_zeros / rem_zeros                  -> remove_zeros(ia, ja, a)
    zmumps(job=6)                           -> solve(rhs, ...)

Every call goes to libmovfem_gpu.so; there is no CPU fallback.  Errors raise
MvfError carrying the library's code and message.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .model import MTModel


class MvfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{L.ERRORS.get(code, code)}: {msg}")
        self.code = code


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def solver_opts(method="cocg", precision="fp64", matrix="c64ref", precond="jacobi", tol=1e-8, maxit=100000,
                check_every=25, inner_tol=1e-2, gmres_m=0, freqs_per_factor=1, direct_tile=0,
                use_graph=False) -> L.SolverOpts:
    o = L.SolverOpts()
    L.lib().mvf_default_solver_opts(C.byref(o))
    o.method = {"cocg": L.MVF_KRYLOV_COCG, "bicgstab": L.MVF_KRYLOV_BICGSTAB, "gmres": L.MVF_KRYLOV_GMRES}[method]
    o.precision = {"fp64": L.MVF_PREC_FP64, "mixed": L.MVF_PREC_MIXED}[precision]
    o.matrix = {"c64ref": L.MVF_MATRIX_C64REF, "fp64": L.MVF_MATRIX_FP64}[matrix]
    o.precond = {"jacobi": L.MVF_PC_JACOBI, "none": L.MVF_PC_NONE, "direct": L.MVF_PC_DIRECT}[precond]
    o.gmres_m = gmres_m
    o.freqs_per_factor = freqs_per_factor
    o.direct_tile = direct_tile
    o.use_graph = 2 if use_graph == "persistent" else int(bool(use_graph))
    o.tol = tol
    o.maxit = maxit
    o.check_every = check_every
    o.inner_tol = inner_tol
    return o


def info_dict(info: L.SolveInfo, ncol: int) -> dict:
    return {"iterations": info.iterations, "refinements": info.refinements, "converged": bool(info.converged),
            "fallbacks": info.fallbacks, "true_resid": [info.true_resid[i] for i in range(min(ncol, 64))],
            "seconds": info.seconds, "spmv_launches": list(info.spmv_launches), "spmv_ms": list(info.spmv_ms),
            "spmv_bytes": list(info.spmv_bytes), "spmv_ncol": info.spmv_ncol,
            "factor_seconds": info.factor_seconds, "factor_flops": info.factor_flops,
            "factor_bytes": info.factor_bytes, "factorizations": info.factorizations,
            "static_pivots": info.static_pivots, "precond_applies": info.precond_applies,
            "gemm_ms": info.gemm_ms, "gemm_flops": info.gemm_flops, "gemm_launches": info.gemm_launches,
            "operator_fp32": bool(info.operator_fp32), "spmv_lpr": info.spmv_lpr,
            "graph_launches": info.graph_launches}


class GpuVfem:
    """One model resident on one GPU (the handle owns every device buffer)."""

    def __init__(self, m: MTModel, emulate: int = L.MVF_EMU_ALL, device: int = 0, verbose: int = 0):
        m.validate()
        self.m = m
        self._lib = L.lib()
        self._keep = [np.ascontiguousarray(m.xp, np.float64), np.ascontiguousarray(m.yp, np.float64),
                      np.ascontiguousarray(m.zp, np.float64),
                      np.ascontiguousarray(m.sigma.astype(np.complex128)).view(np.float64),
                      np.ascontiguousarray(m.mu, np.float64)]
        xp, yp, zp, sg, mu = self._keep
        self._grid = L.Grid(m.nx, m.ny, m.nz, m.nodes_per_elem, _dp(xp), _dp(yp), _dp(zp))
        self._model = L.Model(_dp(sg), _dp(mu))
        self._lsig = np.ascontiguousarray(m.l_sigma, np.float64)
        self._ldz = np.ascontiguousarray(np.append(np.asarray(m.l_dz, np.float64), 0.0))
        self._bc = L.BC(int(m.dirichlet), m.inimod, m.gpml_sch, m.nextd, m.nzl_top, 0, m.a0, m.b0, m.nn,
                        float(m.ztop), float(m.h_sigma), len(self._lsig), 0, _dp(self._lsig), _dp(self._ldz))
        self._opts = L.Options(emulate, device, verbose, 0)
        self.h = C.c_void_p()
        rc = self._lib.mvf_init(C.byref(self._grid), C.byref(self._model), C.byref(self._bc), C.byref(self._opts),
                                C.byref(self.h))
        if rc:
            msg = self._lib.mvf_last_error(self.h).decode() if self.h else ""
            self.close()
            raise MvfError(rc, msg)
        nne = C.c_int32()
        nnze = C.c_int64()
        ntriu = C.c_int64()
        self._check(self._lib.mvf_sizes(self.h, C.byref(nne), C.byref(nnze), C.byref(ntriu)))
        self.nne, self.nnze, self.nnz_triu = nne.value, nnze.value, ntriu.value
        self.me = m.edges_per_elem
        self.omega = None

    def _check(self, rc):
        if rc:
            raise MvfError(rc, self._lib.mvf_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self._lib.mvf_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup_times(self):
        t = np.zeros(4)
        self._check(self._lib.mvf_setup_times(self.h, _dp(t)))
        return {"numbering": t[0], "pattern": t[1], "integration": t[2], "assembly": t[3]}

    def assemble(self, precision: str = "fp64") -> float:
        """Re-integrate and re-assemble K, M from the resident model; returns the
        element kernel's time in ms."""
        ms = np.zeros(1)
        self._check(self._lib.mvf_assemble(self.h, {"fp64": 0, "fp32": 1}[precision], _dp(ms)))
        return float(ms[0])

    def operators(self):
        """(rowptr, col, K, M): the device's full CSR pattern (0-based) and the
        FP64 values of K and M as currently assembled."""
        rp = np.empty(self.nne + 1, np.int64)
        col = np.empty(self.nnze, np.int32)
        K = np.empty(self.nnze)
        Mv = np.empty(self.nnze)
        self._check(self._lib.mvf_get_operators(self.h, rp.ctypes.data_as(C.POINTER(C.c_int64)), _ip(col), _dp(K),
                                                _dp(Mv)))
        return rp, col, K, Mv

    def gne(self) -> np.ndarray:
        out = np.empty(self.m.ne * self.me, np.int32)
        self._check(self._lib.mvf_get_gne(self.h, _ip(out)))
        return out.reshape(self.m.ne, self.me)

    def set_frequency(self, i: int):
        """update_omega(i+1), update_sigma, pset_pmodel for frequency index i."""
        self.set_frequency_hz(float(self.m.freqs[i]))

    def set_frequency_hz(self, f: float):
        self._check(self._lib.mvf_set_frequency(self.h, float(f)))
        self.omega = 2.0 * 3.1415926535897932384626433 * f

    def integrate(self, e0: int, e1: int, want_b: bool = True):
        """K, M upper-packed (n, E(E+1)/2) and b (n, E, 2) for elements [e0, e1)."""
        n = e1 - e0
        np_ = self.me * (self.me + 1) // 2
        K = np.empty((n, np_))
        M = np.empty((n, np_))
        b = np.empty((n, self.me, 2), np.complex128) if want_b else None
        bp = _dp(b.view(np.float64)) if want_b else None
        self._check(self._lib.mvf_integrate(self.h, e0, e1, _dp(K), _dp(M), bp))
        return K, M, b

    def alocal(self, K_packed, M_packed, omega=None):
        """Full (E, E) element matrix K + i fl32(omega) M from packed rows."""
        w = float(np.float32(self.omega if omega is None else omega))
        me = self.me
        iu = np.triu_indices(me)
        A = np.zeros((me, me), np.complex128)
        A[iu] = K_packed + 1j * (w * M_packed)
        A = A + np.triu(A, 1).T
        return A

    def global_vfem(self, out=None):
        """(ia, ja, a, b) exactly as the reference's global_vfem leaves them.
        out: the four arrays of an earlier call to fill again (as a caller that
        allocates them once before its frequency loop does)."""
        if out is None:
            out = (np.empty(self.nnze, np.int32), np.empty(self.nnze, np.int32), np.empty(self.nnze, np.complex128),
                   np.empty(2 * self.nne, np.complex128))
        ia, ja, a, b = out
        for v, dt, n in ((ia, np.int32, self.nnze), (ja, np.int32, self.nnze), (a, np.complex128, self.nnze),
                         (b, np.complex128, 2 * self.nne)):
            if v.dtype != dt or v.size != n or not v.flags.c_contiguous:
                raise ValueError("global_vfem: out arrays must be contiguous int32[nnze] x2, complex128[nnze], "
                                 "complex128[2*nne]")
        self._check(self._lib.mvf_global_vfem(self.h, _ip(ia), _ip(ja), _dp(a.view(np.float64)),
                                               _dp(b.view(np.float64)), self.nnze, self.nne, 2))
        return ia, ja, a, b

    def global_vfem_nz(self):
        """global_vfem followed by find_zeros/rem_zeros, formatted on the
        device: (ia, ja, a, b) with the (0, 0) entries dropped."""
        cap = self.nnz_triu
        ia = np.empty(cap, np.int32)
        ja = np.empty(cap, np.int32)
        a = np.empty(cap, np.complex128)
        b = np.empty(2 * self.nne, np.complex128)
        n = C.c_int64()
        self._check(self._lib.mvf_global_vfem_nz(self.h, _ip(ia), _ip(ja), _dp(a.view(np.float64)),
                                                  _dp(b.view(np.float64)), cap, self.nne, 2, C.byref(n)))
        return ia[:n.value], ja[:n.value], a[:n.value], b

    def remove_zeros(self, ia, ja, a):
        ia, ja = ia.copy(), ja.copy()
        a = np.ascontiguousarray(a, np.complex128).copy()
        n = C.c_int64()
        self._check(self._lib.mvf_remove_zeros(len(ia), _ip(ia), _ip(ja), _dp(a.view(np.float64)), C.byref(n)))
        return ia[:n.value], ja[:n.value], a[:n.value]

    def solve(self, rhs, **kw):
        """zmumps(job=6)-shaped: rhs (2*nne complex, polarization-major) -> solution, info."""
        x = np.ascontiguousarray(rhs, np.complex128).copy()
        o = solver_opts(**kw)
        info = L.SolveInfo()
        rc = self._lib.mvf_solve(self.h, _dp(x.view(np.float64)), C.byref(o), C.byref(info))
        d = info_dict(info, 2)
        if rc:
            raise MvfError(rc, self._lib.mvf_last_error(self.h).decode() + f" {d}")
        return x, d

    def solve_noalloc(self, rhs, **kw):
        """mvf_solve_noalloc: as solve() for COCG/BiCGStab (FP64, Jacobi or no
        preconditioner) with every device buffer in the library's static
        scratch array; no device allocation."""
        x = np.ascontiguousarray(rhs, np.complex128).copy()
        o = solver_opts(**kw)
        info = L.SolveInfo()
        rc = self._lib.mvf_solve_noalloc(self.h, _dp(x.view(np.float64)), C.byref(o), C.byref(info))
        d = info_dict(info, 2)
        if rc:
            raise MvfError(rc, self._lib.mvf_last_error(self.h).decode() + f" {d}")
        return x, d

    def solve_frequencies(self, freqs, **kw):
        """Device RHS for each frequency (in order) and one batched solve."""
        f = np.ascontiguousarray(freqs, np.float64)
        o = solver_opts(**kw)
        info = L.SolveInfo()
        rc = self._lib.mvf_solve_frequencies(self.h, _dp(f), len(f), C.byref(o), C.byref(info))
        d = info_dict(info, 2 * len(f))
        if rc:
            raise MvfError(rc, self._lib.mvf_last_error(self.h).decode() + f" {d}")
        return d

    def solution(self, ncols):
        x = np.empty(self.nne * ncols, np.complex128)
        self._check(self._lib.mvf_get_solution(self.h, _dp(x.view(np.float64)), ncols))
        return x.reshape(ncols, self.nne)

    def rhs(self, ncols):
        x = np.empty(self.nne * ncols, np.complex128)
        self._check(self._lib.mvf_get_rhs(self.h, _dp(x.view(np.float64)), ncols))
        return x.reshape(ncols, self.nne)

    def readout(self, x=None, f=0, ke_top=0, rx=(), ry=(), full=False):
        """node_solution + rx_data on the device.  x: host solution (2*nne,
        polarization-major) at the current frequency, or None for frequency f
        of the last solve_frequencies.  Returns esol, hsol (2, npt, 3) and, with
        receivers, (E, H (nrx, 2, 3), Z (nrx, 4), rho, phase (nrx, 4)).  full=True
        returns a dict that adds the receivers' heights ("rx_elev") and Z, rho_a,
        phase at every lattice node ("node_z" (npt, 4), "node_rho", "node_phi")."""
        npt = self.m.npt
        es = np.empty(2 * npt * 3, np.complex128)
        hs = np.empty(2 * npt * 3, np.complex128)
        rx = np.ascontiguousarray(rx, np.float64)
        ry = np.ascontiguousarray(ry, np.float64)
        n = len(rx)
        oe = np.empty(max(n, 1) * 6, np.complex128)
        oh = np.empty(max(n, 1) * 6, np.complex128)
        oz = np.empty(max(n, 1) * 4, np.complex128)
        rho = np.empty(max(n, 1) * 4)
        phi = np.empty(max(n, 1) * 4)
        xv = None if x is None else np.ascontiguousarray(x, np.complex128)
        elev = np.empty(max(n, 1))
        nz = np.empty(npt * 4, np.complex128) if full else None
        nrho = np.empty(npt * 4) if full else None
        nphi = np.empty(npt * 4) if full else None
        self._check(self._lib.mvf_readout(self.h, None if xv is None else _dp(xv.view(np.float64)), f, ke_top, n,
                                          _dp(rx) if n else None, _dp(ry) if n else None,
                                          _dp(es.view(np.float64)), _dp(hs.view(np.float64)),
                                          _dp(oe.view(np.float64)), _dp(oh.view(np.float64)),
                                          _dp(oz.view(np.float64)), _dp(rho), _dp(phi), _dp(elev),
                                          _dp(nz.view(np.float64)) if full else None,
                                          _dp(nrho) if full else None, _dp(nphi) if full else None))
        if full:
            d = {"esol": es.reshape(2, npt, 3), "hsol": hs.reshape(2, npt, 3), "node_z": nz.reshape(npt, 4),
                 "node_rho": nrho.reshape(npt, 4), "node_phi": nphi.reshape(npt, 4)}
            if n:
                d.update(rx_e=oe[:6 * n].reshape(n, 2, 3), rx_h=oh[:6 * n].reshape(n, 2, 3), rx_z=oz[:4 * n].reshape(n, 4),
                         rx_rho=rho[:4 * n].reshape(n, 4), rx_phi=phi[:4 * n].reshape(n, 4), rx_elev=elev[:n].copy())
            return d
        out = (es.reshape(2, npt, 3), hs.reshape(2, npt, 3))
        if n:
            out += (oe[:6 * n].reshape(n, 2, 3), oh[:6 * n].reshape(n, 2, 3), oz[:4 * n].reshape(n, 4),
                    rho[:4 * n].reshape(n, 4), phi[:4 * n].reshape(n, 4))
        return out


    def problem_fields(self):
        """PROBLEM.OUT's fields at the current frequency (write_problem_elem,
        problem.f90:459-507): primary E, primary H and the source term, each
        complex (2, npt, 3), from the last element holding each node."""
        npt = self.m.npt
        out = [np.empty(2 * npt * 3, np.complex128) for _ in range(3)]
        self._check(self._lib.mvf_problem_fields(self.h, *(_dp(a.view(np.float64)) for a in out)))
        return tuple(a.reshape(2, npt, 3) for a in out)


def solve_batch(handles, freqs, nthreads=0, **kw):
    """mvf_solve_batch: every handle (model) solves its row of freqs (nh, nfreq)
    with solve_frequencies' semantics, several models in flight at once.
    Returns (infos, status codes)."""
    hs = (C.c_void_p * len(handles))(*[g.h.value for g in handles])
    f = np.ascontiguousarray(freqs, np.float64)
    if f.ndim == 1:
        f = np.tile(f, (len(handles), 1))
    nh, nf = f.shape
    o = solver_opts(**kw)
    infos = (L.SolveInfo * nh)()
    st = np.zeros(nh, np.int32)
    rc = L.lib().mvf_solve_batch(hs, nh, _dp(f), nf, C.byref(o), infos, _ip(st), int(nthreads))
    out = [info_dict(infos[i], 2 * nf) for i in range(nh)]
    if rc:
        bad = [i for i in range(nh) if st[i]]
        raise MvfError(rc, f"models {bad} failed: " + "; ".join(
            f"{i}: {handles[i]._lib.mvf_last_error(handles[i].h).decode()} (iterations {out[i]['iterations']}, "
            f"factorizations {out[i]['factorizations']}, static pivots {out[i]['static_pivots']}, true residual "
            f"{max(out[i]['true_resid'][:2 * nf]):.2e})" for i in bad))
    return out, st


# ---------------------------------------------------------------- text output
def surface_nodes(m: MTModel, nzl_top: int, nzl_next: int):
    """The surface points SURFACE.OUT lists (sur_data, receiver_data.f90:145-175):
    lattice nodes i, j inside the nextd extension cells on the lattice layer
    nnz - (nzl_top + nzl_next)*(nord-1), y outer and x inner.  Returns nxb, nyb,
    xs (nxb), ys (nyb), zs (nsp), 0-based node ids (nsp)."""
    o = m.nord - 1
    i1, i2 = m.nextd * o, m.nnx - m.nextd * o - 1
    j1, j2 = m.nextd * o, m.nny - m.nextd * o - 1
    k = m.nnz - nzl_top * o - nzl_next * o - 1
    if not (0 <= i1 <= i2 and 0 <= j1 <= j2 and 0 <= k < m.nnz):
        raise ValueError("surface selection does not fit the lattice")
    ii, jj = np.meshgrid(np.arange(i1, i2 + 1), np.arange(j1, j2 + 1), indexing="xy")
    ids = (ii.ravel() * m.nny + jj.ravel()) * m.nnz + k
    return (i2 - i1 + 1, j2 - j1 + 1, np.ascontiguousarray(m.xp[i1:i2 + 1]), np.ascontiguousarray(m.yp[j1:j2 + 1]),
            np.ascontiguousarray(m.zp[ids]), np.ascontiguousarray(ids, np.int64))


def _wcheck(rc, what):
    if rc:
        raise MvfError(rc, f"{what} failed")


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dt)


def write_solution(path, nf, esol, hsol, node_z, node_rho, node_phi, sel=None, append=False):
    """SOLUTION.OUT block of frequency nf (write_solution, solution.f90:73-205)."""
    es, hs, nz = _c(esol, np.complex128), _c(hsol, np.complex128), _c(node_z, np.complex128)
    nr, npf = _c(node_rho), _c(node_phi)
    npt = es.size // 6
    s = None if sel is None else _c(sel, np.int32)
    _wcheck(L.lib().mvf_write_solution(str(path).encode(), nf, int(append), npt, _dp(es.view(np.float64)),
                                       _dp(hs.view(np.float64)), _dp(nz.view(np.float64)), _dp(nr), _dp(npf),
                                       0 if s is None else len(s), None if s is None else _ip(s)),
            "mvf_write_solution")


def write_receivers(path, nf, nfreq, rx, ry, elev, rx_e, rx_h, rx_z, rho, phi, append=False):
    """RECEIVER.OUT block of frequency nf (write_receiver_data, receiver_data.f90:25-67)."""
    rx, ry, el = _c(rx), _c(ry), _c(elev)
    e, h, z = _c(rx_e, np.complex128), _c(rx_h, np.complex128), _c(rx_z, np.complex128)
    r, p = _c(rho), _c(phi)
    _wcheck(L.lib().mvf_write_receivers(str(path).encode(), nf, int(append), nfreq, len(rx), _dp(rx), _dp(ry),
                                        _dp(el), _dp(e.view(np.float64)), _dp(h.view(np.float64)),
                                        _dp(z.view(np.float64)), _dp(r), _dp(p)), "mvf_write_receivers")


def write_surface(path, nf, nfreq, surface, esol, hsol, node_z, node_rho, node_phi, append=False):
    """SURFACE.OUT block of frequency nf (write_surface_data, receiver_data.f90:69-109);
    surface = surface_nodes(...)."""
    nxb, nyb, xs, ys, zs, ids = surface
    es, hs, nz = _c(esol, np.complex128), _c(hsol, np.complex128), _c(node_z, np.complex128)
    nr, npf = _c(node_rho), _c(node_phi)
    _wcheck(L.lib().mvf_write_surface(str(path).encode(), nf, int(append), nfreq, nxb, nyb, _dp(_c(xs)), _dp(_c(ys)),
                                      _dp(_c(zs)), _c(ids, np.int64).ctypes.data_as(C.POINTER(C.c_int64)),
                                      es.size // 6, _dp(es.view(np.float64)), _dp(hs.view(np.float64)),
                                      _dp(nz.view(np.float64)), _dp(nr), _dp(npf)), "mvf_write_surface")


def write_problem(path, nf, ep, hp, src, sel=None, append=False):
    """PROBLEM.OUT block of frequency nf (write_problem_ugrid, problem.f90:188-240)."""
    e, h, s = (_c(a, np.complex128) for a in (ep, hp, src))
    npt = e.size // 6
    q = None if sel is None else _c(sel, np.int32)
    _wcheck(L.lib().mvf_write_problem(str(path).encode(), nf, int(append), npt, _dp(e.view(np.float64)),
                                      _dp(h.view(np.float64)), _dp(s.view(np.float64)),
                                      0 if q is None else len(q), None if q is None else _ip(q)),
            "mvf_write_problem")
