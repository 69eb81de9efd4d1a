"""Input files and lattice generation: the reference's input() and PARAM.INP
reads (read_input.f90:37-201, MoVFEM_3DMT.f90:302-334) and grid_3d
(geometry.f90:48-122), run by libmovfem_gpu.so's host code.

    call input(); read PARAM.INP      -> read_inputs(inp, param)
    call grid_3d(in_dx, ..., nf, f)   -> grid_3d(mesh_input)
    initialize_vfem's module state    -> to_model(mesh, params), receivers(mesh)

Array layouts follow the reference: interface points as (nsf, npoints) with
the input interfaces in slots 1..nsf-3 (0-based; 0, nsf-2 and nsf-1 are the
generated bottom, air and top slots), the inner model's heights zm indexed
(i, j, k) with k fastest and its conductivity as (n_points, isigma).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from . import _lib as L
from .model import MTModel
from .vfem import MvfError


def _err(rc, what):
    if rc:
        msg = L.lib().mvf_last_error(None)
        raise MvfError(rc, f"{what}: {msg.decode() if msg else ''}")


def _arr(ptr, n, dt=np.float64):
    if n == 0 or not ptr:
        return np.zeros(0, dt)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)


@dataclasses.dataclass
class MeshInput:
    """grid_3d's arguments (geometry.f90:48-62)."""

    dx: float
    dy: float
    dz: float
    high: float
    depth: float
    sigma_bg: float
    mn: int
    xsr: np.ndarray
    ysr: np.ndarray
    zsr: np.ndarray
    nsp: np.ndarray       # (nsf,) int32
    xto: np.ndarray       # (nsf, npoints)
    yto: np.ndarray
    zto: np.ndarray
    xm: np.ndarray        # (mx,)
    ym: np.ndarray        # (my,)
    zm: np.ndarray        # (mx*my*mz,), k fastest
    mz: int
    ijsigma: np.ndarray   # (isigma, 2)
    sigma: np.ndarray     # (mx*my*mz, isigma)
    ijmu: np.ndarray | None = None  # (imu, 2) or None for mu_r = 1
    mu: np.ndarray | None = None    # (mx*my*mz, imu)
    freqs: np.ndarray | None = None

    @property
    def nsf(self) -> int:
        return len(self.nsp)

    @classmethod
    def build(cls, dx, dy, dz, high, depth, sigma_bg, mn, sites, interfaces, xm, ym, zm, sigma, ijsigma=((1, 1),),
              mu=None, ijmu=None, freqs=None) -> "MeshInput":
        """From plain lists: sites (nsr, 3) x, y, z; interfaces a list of
        (n, 3) point sets, deepest first; zm (mx, my, mz); sigma
        (mx, my, mz, isigma) or (mx, my, mz) for one component."""
        sites = np.asarray(sites, np.float64).reshape(-1, 3)
        nsf = len(interfaces) + 3
        nmax = max(len(p) for p in interfaces)
        xto, yto, zto = (np.zeros((nsf, nmax)) for _ in range(3))
        nsp = np.zeros(nsf, np.int32)
        for l, p in enumerate(interfaces, start=1):
            p = np.asarray(p, np.float64).reshape(-1, 3)
            nsp[l] = len(p)
            xto[l, :len(p)], yto[l, :len(p)], zto[l, :len(p)] = p[:, 0], p[:, 1], p[:, 2]
        zm = np.asarray(zm, np.float64)
        mx, my, mz = zm.shape
        sg = np.asarray(sigma, np.float64).reshape(mx * my * mz, -1)
        m = None if mu is None else np.asarray(mu, np.float64).reshape(mx * my * mz, -1)
        return cls(float(dx), float(dy), float(dz), float(high), float(depth), float(sigma_bg), int(mn),
                   sites[:, 0].copy(), sites[:, 1].copy(), sites[:, 2].copy(), nsp, xto, yto, zto,
                   np.asarray(xm, np.float64), np.asarray(ym, np.float64), zm.ravel(), mz,
                   np.asarray(ijsigma, np.int32).reshape(-1, 2), sg,
                   None if ijmu is None else np.asarray(ijmu, np.int32).reshape(-1, 2), m,
                   None if freqs is None else np.asarray(freqs, np.float64))


@dataclasses.dataclass
class RunParams:
    """PARAM.INP (MoVFEM_3DMT.f90:302-334)."""

    freqs: np.ndarray
    boundary: int
    inimod: int
    h_sigma: float
    l_sigma: np.ndarray
    l_dz: np.ndarray
    gpml_sch: int = 0
    a0: float = 0.0
    b0: float = 0.0
    nn: float = 0.0


@dataclasses.dataclass
class Mesh:
    """grid_3d's output: the lattice and model (geometry.f90:17-26)."""

    nx: int
    ny: int
    nz: int
    mn: int
    nextd: int
    ztop: float
    x0: float
    y0: float
    z0: float
    omega: float
    skin_depth: float
    xp: np.ndarray
    yp: np.ndarray
    zp: np.ndarray
    sigma: np.ndarray   # (npt, 6) complex
    mu: np.ndarray      # (npt, 6)
    nzl: np.ndarray     # (nsf,)
    xsr: np.ndarray     # shifted site coordinates
    ysr: np.ndarray
    zsr: np.ndarray


def read_inputs(inp_path=None, param_path=None):
    """(MeshInput or None, RunParams or None) from 3DEM_FEM.INP / PARAM.INP."""
    lib = L.lib()
    h = C.c_void_p()
    enc = (lambda p: None if p is None else os.fspath(p).encode())
    _err(lib.mvf_read_inputs(enc(inp_path), enc(param_path), C.byref(h)), "mvf_read_inputs")
    try:
        mi, pr = L.MeshInput(), L.Params()
        _err(lib.mvf_inputs_view(h, C.byref(mi), C.byref(pr)), "mvf_inputs_view")
        mesh_in = params = None
        if param_path is not None:
            params = RunParams(_arr(pr.freqs, pr.nf), pr.boundary, pr.inimod, pr.h_sigma, _arr(pr.l_sigma, pr.nl),
                               _arr(pr.l_dz, pr.nl - 1), pr.gpml_sch, pr.a0, pr.b0, pr.nn)
        if inp_path is not None:
            nsf, npts = mi.nsf, mi.mx * mi.my * mi.mz
            nsp = _arr(mi.nsp, nsf, np.int32)
            nmax = int(nsp.max()) if nsf else 0
            to = [_arr(getattr(mi, k), nsf * nmax).reshape(nmax, nsf).T.copy() for k in ("xto", "yto", "zto")]
            ijs = _arr(mi.ijsigma, 2 * mi.isigma, np.int32).reshape(2, mi.isigma).T.copy()
            ijm = None if mi.imu == 0 else _arr(mi.ijmu, 2 * mi.imu, np.int32).reshape(2, mi.imu).T.copy()
            mu = None if mi.imu == 0 else _arr(mi.mu, mi.imu * npts).reshape(npts, mi.imu)
            mesh_in = MeshInput(mi.dx, mi.dy, mi.dz, mi.high, mi.depth, mi.sigma_bg, mi.mn, _arr(mi.xsr, mi.nsr),
                                _arr(mi.ysr, mi.nsr), _arr(mi.zsr, mi.nsr), nsp, *to, _arr(mi.xm, mi.mx),
                                _arr(mi.ym, mi.my), _arr(mi.zm, npts), mi.mz, ijs,
                                _arr(mi.sigma, mi.isigma * npts).reshape(npts, mi.isigma), ijm, mu,
                                None if params is None else params.freqs.copy())
    finally:
        lib.mvf_inputs_free(h)
    return mesh_in, params


def grid_3d(mi: MeshInput, freqs=None) -> Mesh:
    """The lattice and model grid_3d builds from mi (frequencies: freqs, else
    mi.freqs)."""
    lib = L.lib()
    fr = np.ascontiguousarray(mi.freqs if freqs is None else freqs, np.float64)
    if fr.ndim != 1 or len(fr) == 0:
        raise ValueError("at least one frequency is needed")
    d = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
    keep = dict(xsr=d(mi.xsr), ysr=d(mi.ysr), zsr=d(mi.zsr), nsp=np.ascontiguousarray(mi.nsp, np.int32),
                xto=np.asfortranarray(mi.xto, np.float64), yto=np.asfortranarray(mi.yto, np.float64),
                zto=np.asfortranarray(mi.zto, np.float64), xm=d(mi.xm), ym=d(mi.ym), zm=d(mi.zm),
                ijsigma=np.asfortranarray(mi.ijsigma, np.int32), sigma=d(mi.sigma), fr=fr)
    s = L.MeshInput()
    s.dx, s.dy, s.dz, s.high, s.depth, s.sigma_bg = mi.dx, mi.dy, mi.dz, mi.high, mi.depth, mi.sigma_bg
    s.mn, s.nsr, s.nsf = mi.mn, len(keep["xsr"]), mi.nsf
    s.mx, s.my, s.mz = len(keep["xm"]), len(keep["ym"]), mi.mz
    if len(keep["zm"]) != s.mx * s.my * s.mz:
        raise ValueError("zm must hold mx*my*mz heights")
    s.isigma = keep["ijsigma"].shape[0]
    if keep["sigma"].size != s.isigma * len(keep["zm"]):
        raise ValueError("sigma must be (mx*my*mz, isigma)")
    s.imu = 0 if mi.ijmu is None else len(mi.ijmu)
    if s.imu:
        keep["ijmu"] = np.asfortranarray(mi.ijmu, np.int32)
        keep["mu"] = d(mi.mu)
        if keep["mu"].size != s.imu * len(keep["zm"]):
            raise ValueError("mu must be (mx*my*mz, imu)")
        s.ijmu = keep["ijmu"].ctypes.data_as(L._ip)
        s.mu = keep["mu"].ctypes.data_as(L._dp)
    if keep["xto"].shape[0] != mi.nsf:
        raise ValueError("interface arrays must be (nsf, npoints)")
    s.nf = len(fr)
    for k in ("xsr", "ysr", "zsr", "xto", "yto", "zto", "xm", "ym", "zm", "sigma"):
        setattr(s, k, keep[k].ctypes.data_as(L._dp))
    s.nsp = keep["nsp"].ctypes.data_as(L._ip)
    s.ijsigma = keep["ijsigma"].ctypes.data_as(L._ip)
    s.freqs = fr.ctypes.data_as(L._dp)
    h = C.c_void_p()
    _err(lib.mvf_grid3d(C.byref(s), C.byref(h)), "mvf_grid3d")
    try:
        dims = np.zeros(12, np.int64)
        vals = np.zeros(6)
        _err(lib.mvf_mesh_sizes(h, dims.ctypes.data_as(C.POINTER(C.c_int64)), vals.ctypes.data_as(L._dp)),
             "mvf_mesh_sizes")
        nx, ny, nz, mn, nnx, nny, nnz, npt, nextd, nsf, nsr, _ = (int(v) for v in dims)
        xp, yp, zp = np.zeros(nnx), np.zeros(nny), np.zeros(npt)
        sg, mu = np.zeros((npt, 6), np.complex128), np.zeros((npt, 6))
        nzl = np.zeros(nsf, np.int32)
        xs, ys, zs = np.zeros(nsr), np.zeros(nsr), np.zeros(nsr)
        _err(lib.mvf_mesh_arrays(h, *(a.ctypes.data_as(L._dp) for a in (xp, yp, zp)),
                                 sg.view(np.float64).ctypes.data_as(L._dp), mu.ctypes.data_as(L._dp),
                                 nzl.ctypes.data_as(L._ip), *(a.ctypes.data_as(L._dp) for a in (xs, ys, zs))),
             "mvf_mesh_arrays")
    finally:
        lib.mvf_mesh_free(h)
    return Mesh(nx, ny, nz, mn, nextd, *(float(v) for v in vals), xp, yp, zp, sg, mu, nzl, xs, ys, zs)


def to_model(mesh: Mesh, params: RunParams | None = None, freqs=None, name: str = "input") -> MTModel:
    """The module state initialize_vfem hands the hot path
    (MoVFEM_3DMT.f90:337-388): lattice, model, boundary settings."""
    p = params
    fr = np.asarray(freqs if freqs is not None else (p.freqs if p is not None else [mesh.omega / (2 * np.pi)]),
                    np.float64)
    m = MTModel(mesh.nx, mesh.ny, mesh.nz, mesh.mn, mesh.xp, mesh.yp, mesh.zp, mesh.sigma, mesh.mu, fr,
                nextd=mesh.nextd, nzl_top=int(mesh.nzl[-1]), ztop=mesh.ztop, name=name)
    if p is not None:
        m.dirichlet = p.boundary == 1
        m.inimod = p.inimod
        m.h_sigma = p.h_sigma
        m.l_sigma = tuple(float(v) for v in p.l_sigma)
        m.l_dz = tuple(float(v) for v in p.l_dz)
        if p.boundary == 0:
            m.gpml_sch, m.a0, m.b0, m.nn = p.gpml_sch, p.a0, p.b0, p.nn
    m.validate()
    return m


def receivers(mesh: Mesh):
    """Receiver sites on the earth's surface (rx_data, receiver_data.f90:194-230):
    the shifted site coordinates and the top earth element layer ke_top
    (g_nz - g_nzl(g_nsf) - g_nzl(g_nsf-1) - 1)."""
    return {"x": mesh.xsr.copy(), "y": mesh.ysr.copy(),
            "ke_top": mesh.nz - int(mesh.nzl[-1]) - int(mesh.nzl[-2]),
            "nzl_top": int(mesh.nzl[-1]), "nzl_next": int(mesh.nzl[-2])}


def load_case(directory):
    """3DEM_FEM.INP and PARAM.INP of a run directory -> (MTModel, Mesh, RunParams)."""
    d = os.fspath(directory)
    mi, p = read_inputs(os.path.join(d, "3DEM_FEM.INP"), os.path.join(d, "PARAM.INP"))
    mesh = grid_3d(mi, p.freqs)
    return to_model(mesh, p, name=os.path.basename(os.path.abspath(d))), mesh, p
