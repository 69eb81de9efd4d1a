"""ctypes front end of the CPU oracle (movfem_oracle.c) and of the reference build.

TEST INFRASTRUCTURE ONLY -- the checker the HIP path is compared against.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.

Pinning status (what each restated function is checked against):
  gne numbering (c_gne12/36/54)          -- reference build (_ref) on tiny grids, golden fixtures,
                                            independent geometric derivation, DOF-count formulas
  slot map (ga_nzindx/shr_nzindx*)       -- reference build, golden fixtures, sort-rule invariants
  element matrices (alocal)              -- reference build, golden fixtures, closed-form brick
                                            matrices (8-node), curl-curl null space K.G = 0
  element RHS (blocal, p_source)         -- reference build, golden fixtures, closed form of the
                                            primary field on a box
  global_vfem (assembly + sort, c64)     -- reference build, golden fixtures
  update_sigma linear-index semantics    -- reference build
  Dirichlet values (f_boundary, p_pfields
    of the boundary module, Wait recursion,
    ezl) and their lift, inimod 2/3         -- reference build (per-element values, global_vfem RHS,
                                            node_solution), closed forms of the half-space and
                                            one-layer cases, tangential-trace identity on a box
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(str(path))
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int] * 4 + [dp] * 5 + [C.c_int] * 3 + [C.c_double] * 3 + [C.c_int] * 2
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_set_omega.argtypes = [C.c_void_p, C.c_double]
        L.orc_update_sigma.argtypes = [C.c_void_p]
        L.orc_get_sigma.argtypes = [C.c_void_p, dp]
        L.orc_build_gne.restype = C.c_int32
        L.orc_build_gne.argtypes = [C.c_void_p]
        L.orc_build_nzindx.restype = C.c_int64
        L.orc_build_nzindx.argtypes = [C.c_void_p]
        L.orc_get_gne.argtypes = [C.c_void_p, ip]
        L.orc_get_nzindx.argtypes = [C.c_void_p, ip]
        L.orc_element.restype = C.c_int
        L.orc_element.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_set_in_pml.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_global_vfem.restype = C.c_int
        L.orc_global_vfem.argtypes = [C.c_void_p, ip, ip, dp, dp]
        L.orc_remove_zeros.restype = C.c_int64
        L.orc_remove_zeros.argtypes = [C.c_int64, ip, ip, dp]
        L.orc_node_solution.restype = C.c_int
        L.orc_node_solution.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_problem_fields.restype = C.c_int
        L.orc_problem_fields.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_receivers.restype = C.c_int
        L.orc_receivers.argtypes = [C.c_void_p, dp, dp, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp]
        L.orc_set_bdmodel.restype = C.c_int
        L.orc_set_bdmodel.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, dp, dp]
        L.orc_dirichlet_values.restype = C.c_int
        L.orc_dirichlet_values.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, dp]
        L.orc_bd_field.restype = None
        L.orc_bd_field.argtypes = [C.c_void_p, C.c_double, dp]
        L.orc_impedance.restype = None
        L.orc_impedance.argtypes = [dp, dp, C.c_double, dp, dp, dp]
        L.orc_element_loop_timed.restype = C.c_int
        L.orc_element_loop_timed.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Oracle:
    """CPU restatement bound to one model; mirrors the reference's call order:
    construct (initialize_vfem), set_frequency(i) (update_omega, update_sigma,
    pset_pmodel), global_vfem()."""

    def __init__(self, m):
        self.m = m
        L = lib()
        self._keep = [np.ascontiguousarray(m.xp, np.float64), np.ascontiguousarray(m.yp, np.float64),
                      np.ascontiguousarray(m.zp, np.float64),
                      np.ascontiguousarray(m.sigma.astype(np.complex128)).view(np.float64),
                      np.ascontiguousarray(m.mu, np.float64)]
        xp, yp, zp, sg, mu = self._keep
        self.h = L.orc_create(m.nx, m.ny, m.nz, m.nodes_per_elem, _dp(xp), _dp(yp), _dp(zp), _dp(sg), _dp(mu),
                              int(m.dirichlet), m.inimod, m.gpml_sch, m.a0, m.b0, m.nn, m.nextd, m.nzl_top)
        if not self.h:
            raise ValueError("oracle rejected the model")
        ls = np.ascontiguousarray(m.l_sigma, np.float64)
        ld = np.ascontiguousarray(np.append(np.asarray(m.l_dz, np.float64), 0.0))
        if L.orc_set_bdmodel(self.h, float(m.ztop), float(m.h_sigma), len(ls), _dp(ls), _dp(ld)) != 0:
            raise ValueError("oracle rejected the boundary model")
        L.orc_set_omega(self.h, m.omega(0))
        self.omega = m.omega(0)
        self.me = m.edges_per_elem
        self._nne = None
        self._nnze = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().orc_free(self.h)
        except Exception:
            pass

    def set_frequency(self, i: int):
        self.omega = self.m.omega(i)
        L = lib()
        L.orc_set_omega(self.h, self.m.omega(i))
        L.orc_update_sigma(self.h)

    def element_loop_timed(self, e0: int, e1: int, nthreads: int = 1):
        """int_elem_params + alocal (pairs im <= jm) + blocal for elements [e0, e1) in
        reference loop order on nthreads threads; returns (seconds, threads used,
        checksum (sum Re a, sum Im a, sum |b|))."""
        import time

        out = np.zeros(4)
        t0 = time.perf_counter()
        rc = lib().orc_element_loop_timed(self.h, int(e0), int(e1), int(nthreads), _dp(out))
        dt = time.perf_counter() - t0
        if rc:
            raise ValueError(f"oracle element loop failed: {rc}")
        return dt, int(out[3]), out[:3].copy()

    def dirichlet_values(self, ie, je, ke) -> np.ndarray:
        """f_boundary of every edge of element (ie,je,ke) (1-based), complex (me, 2)
        by polarization; zero for interior edges."""
        out = np.zeros(self.me * 4)
        rc = lib().orc_dirichlet_values(self.h, ie, je, ke, _dp(out))
        if rc:
            raise RuntimeError(f"orc_dirichlet_values failed ({rc})")
        return out.view(np.complex128).reshape(self.me, 2)

    def bd_field(self, zn: float) -> np.ndarray:
        """Boundary primary field at node height zn: complex (2,) = (Ex of
        polarization 1, Ey of polarization 2) at the current frequency."""
        out = np.zeros(4)
        lib().orc_bd_field(self.h, float(zn), _dp(out))
        return out.view(np.complex128)

    def sigma(self) -> np.ndarray:
        out = np.empty(self.m.npt * 12)
        lib().orc_get_sigma(self.h, _dp(out))
        return out.view(np.complex128).reshape(self.m.npt, 6)

    def gne(self) -> np.ndarray:
        L = lib()
        self._nne = L.orc_build_gne(self.h)
        out = np.empty(self.m.ne * self.me, np.int32)
        L.orc_get_gne(self.h, _ip(out))
        return out.reshape(self.m.ne, self.me)

    @property
    def nne(self):
        if self._nne is None:
            self.gne()
        return self._nne

    def nzindx(self) -> np.ndarray:
        L = lib()
        if self._nne is None:
            self.gne()
        self._nnze = L.orc_build_nzindx(self.h)
        out = np.empty(self.m.ne * self.me * self.me, np.int32)
        L.orc_get_nzindx(self.h, _ip(out))
        return out.reshape(self.m.ne, self.me, self.me)

    @property
    def nnze(self):
        if self._nnze is None:
            self.nzindx()
        return self._nnze

    def set_in_pml(self, a, b, c):
        lib().orc_set_in_pml(self.h, a, b, c)

    def element(self, ie, je, ke):
        """alocal (me, me) and blocal (me, 2) for the element at 1-based loop
        position (ie, je, ke); advances the GPML flag state like global_vfem."""
        me = self.me
        a = np.empty(me * me * 2)
        b = np.empty(me * 2 * 2)
        rc = lib().orc_element(self.h, ie, je, ke, _dp(a), _dp(b))
        if rc:
            raise RuntimeError(f"orc_element failed ({rc})")
        return a.view(np.complex128).reshape(me, me), b.view(np.complex128).reshape(me, 2)

    def global_vfem(self):
        """(ia, ja, a, b) exactly as the reference's global_vfem leaves them:
        1-based row-major indices of the full pattern, complex64-rounded values
        in column-sorted order, RHS (2*nne) polarization-major."""
        n = self.nnze
        nne = self.nne
        ia = np.empty(n, np.int32)
        ja = np.empty(n, np.int32)
        a = np.empty(2 * n)
        b = np.empty(4 * nne)
        rc = lib().orc_global_vfem(self.h, _ip(ia), _ip(ja), _dp(a), _dp(b))
        if rc:
            raise RuntimeError(f"orc_global_vfem failed ({rc})")
        return ia, ja, a.view(np.complex128), b.view(np.complex128)

    def problem_fields(self):
        """write_problem_elem's values at every lattice node (last element in
        reference order wins): primary E, primary H and the source term, each
        complex (2, npt, 3) for polarizations 1, 2 at the current frequency."""
        npt = self.m.npt
        out = [np.zeros(2 * npt * 3, np.complex128) for _ in range(3)]
        rc = lib().orc_problem_fields(self.h, *(_dp(a.view(np.float64)) for a in out))
        if rc:
            raise RuntimeError(f"orc_problem_fields failed ({rc})")
        return tuple(a.reshape(2, npt, 3) for a in out)

    def node_solution(self, x):
        """esol, hsol (2, npt, 3) complex: total E and H at every lattice node for
        the solution x (2*nne, polarization-major) at the current frequency."""
        npt = self.m.npt
        xv = np.ascontiguousarray(np.asarray(x, np.complex128).ravel()).view(np.float64)
        es = np.empty(2 * npt * 3 * 2)
        hs = np.empty(2 * npt * 3 * 2)
        rc = lib().orc_node_solution(self.h, _dp(xv), _dp(es), _dp(hs))
        if rc:
            raise RuntimeError(f"orc_node_solution failed ({rc})")
        return es.view(np.complex128).reshape(2, npt, 3), hs.view(np.complex128).reshape(2, npt, 3)

    def receivers(self, esol, hsol, ke_top, rx, ry):
        """E, H (nrx, 2, 3), Z (nrx, 4), rho_a, phase (nrx, 4) at receivers."""
        rx = np.ascontiguousarray(rx, np.float64)
        ry = np.ascontiguousarray(ry, np.float64)
        n = len(rx)
        es = np.ascontiguousarray(esol, np.complex128).view(np.float64)
        hs = np.ascontiguousarray(hsol, np.complex128).view(np.float64)
        oe, oh, oz = np.empty(n * 12), np.empty(n * 12), np.empty(n * 8)
        rho, phi = np.empty(n * 4), np.empty(n * 4)
        rc = lib().orc_receivers(self.h, _dp(es), _dp(hs), ke_top, n, _dp(rx), _dp(ry), _dp(oe), _dp(oh), _dp(oz),
                                 _dp(rho), _dp(phi))
        if rc:
            raise RuntimeError(f"orc_receivers failed ({rc})")
        return (oe.view(np.complex128).reshape(n, 2, 3), oh.view(np.complex128).reshape(n, 2, 3),
                oz.view(np.complex128).reshape(n, 4), rho.reshape(n, 4), phi.reshape(n, 4))

    @staticmethod
    def impedance(e, h, omega):
        """Z (4), rho_a (4), phase (4) from e, h = (x pol1, x pol2, y pol1, y pol2)."""
        ev = np.ascontiguousarray(e, np.complex128).view(np.float64)
        hv = np.ascontiguousarray(h, np.complex128).view(np.float64)
        z, rho, phi = np.empty(8), np.empty(4), np.empty(4)
        lib().orc_impedance(_dp(ev), _dp(hv), float(omega), _dp(z), _dp(rho), _dp(phi))
        return z.view(np.complex128), rho, phi

    def node_impedance(self, esol, hsol):
        """z_rho_phi (solution.f90:51-69): Z (npt, 4), rho_a, phase (npt, 4) at every
        lattice node at the current frequency."""
        npt = self.m.npt
        z, rho, phi = np.empty((npt, 4), np.complex128), np.empty((npt, 4)), np.empty((npt, 4))
        es, hs = np.asarray(esol), np.asarray(hsol)
        for i in range(npt):
            e = np.array([es[0, i, 0], es[1, i, 0], es[0, i, 1], es[1, i, 1]])
            h = np.array([hs[0, i, 0], hs[1, i, 0], hs[0, i, 1], hs[1, i, 1]])
            z[i], rho[i], phi[i] = self.impedance(e, h, self.omega)
        return z, rho, phi

    @staticmethod
    def remove_zeros(ia, ja, a):
        ia, ja = ia.copy(), ja.copy()
        av = np.ascontiguousarray(a.astype(np.complex128)).view(np.float64).copy()
        k = lib().orc_remove_zeros(len(ia), _ip(ia), _ip(ja), _dp(av))
        return ia[:k], ja[:k], av.view(np.complex128)[:k]


def ref_available() -> bool:
    return (HERE / "_ref" / "libmovfem_ref.so").exists()
