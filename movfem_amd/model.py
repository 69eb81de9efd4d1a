"""Structured-lattice MT models: the module state the hot path consumes.

The reference keeps this state in module variables filled by its mesh generator
(geometry.f90:17-26): a node lattice g_xp(nnx), g_yp(nny), g_zp(npt) with node
id (i-1)*nny*nnz + (j-1)*nnz + k (z fastest, n_fem.f90:81), a complex
conductivity tensor g_sigma(6, npt) in the order [11,12,13,22,23,33] carrying
+i*float32(omega*eps) on the diagonal (geometry.f90:1049-1064; air nodes hold
only that term, :934-947) and a real permeability tensor g_mu(6, npt).

The mesh generator itself (extension zones, Shepard topography, input files) is
outside the hot path; the builders below produce lattices of the shapes the
benchmark workloads name, with z measured from the bottom of the grid
(z0 = zbot, geometry.f90:342).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

PI = 3.1415926535897932384626433
EPS0 = 8.854187817e-12
MU0 = 4.0 * PI * 1.0e-7

NODES_TO_EDGES = {8: 12, 20: 36, 27: 54}


def f32(x):
    """Round a double to float32 and back (the reference's default-kind CMPLX)."""
    return float(np.float32(x))


@dataclasses.dataclass
class MTModel:
    """One structured-lattice model: grid, materials, boundary settings, frequencies."""

    nx: int  # elements along x (reference g_nx - 1)
    ny: int
    nz: int
    nodes_per_elem: int  # 8, 20 or 27 (reference nf_mn)
    xp: np.ndarray  # (nnx,)
    yp: np.ndarray  # (nny,)
    zp: np.ndarray  # (npt,)  node-indexed z, z fastest
    sigma: np.ndarray  # complex128 (npt, 6)
    mu: np.ndarray  # float64 (npt, 6)
    freqs: np.ndarray  # Hz; freqs[0] is the one baked into sigma's imaginary part
    dirichlet: bool = True
    inimod: int = 1
    gpml_sch: int = 0
    a0: float = 1.0
    b0: float = 1.0
    nn: float = 2.0
    nextd: int = 2
    nzl_top: int = 2
    ztop: float = 0.0
    # boundary 1-D model for Dirichlet inimod 2 (half-space h_sigma) and
    # inimod 3 (layers l_sigma with thicknesses l_dz in km, top layer first)
    h_sigma: float = 0.01
    l_sigma: tuple = (0.01,)
    l_dz: tuple = ()
    name: str = "model"

    @property
    def nord(self) -> int:
        return 2 if self.nodes_per_elem == 8 else 3

    @property
    def edges_per_elem(self) -> int:
        return NODES_TO_EDGES[self.nodes_per_elem]

    @property
    def nnx(self) -> int:
        return self.nx * (self.nord - 1) + 1

    @property
    def nny(self) -> int:
        return self.ny * (self.nord - 1) + 1

    @property
    def nnz(self) -> int:
        return self.nz * (self.nord - 1) + 1

    @property
    def npt(self) -> int:
        return self.nnx * self.nny * self.nnz

    @property
    def ne(self) -> int:
        return self.nx * self.ny * self.nz

    def omega(self, i: int = 0) -> float:
        return 2.0 * PI * float(self.freqs[i])

    def validate(self) -> None:
        if self.nodes_per_elem not in NODES_TO_EDGES:
            raise ValueError("nodes_per_elem must be 8, 20 or 27")
        if min(self.nx, self.ny, self.nz) < 1:
            raise ValueError("grid needs at least one element per direction")
        if self.xp.shape != (self.nnx,) or self.yp.shape != (self.nny,):
            raise ValueError("xp/yp lengths do not match the lattice")
        if self.zp.shape != (self.npt,):
            raise ValueError("zp must have one entry per lattice node")
        if self.sigma.shape != (self.npt, 6) or self.mu.shape != (self.npt, 6):
            raise ValueError("sigma/mu must be (npt, 6)")


def _lattice(n_elem: int, h: float, nord: int) -> np.ndarray:
    sub = nord - 1
    return np.arange(n_elem * sub + 1, dtype=np.float64) * (h / sub)


def node_grid(m: MTModel):
    """Lattice indices (i, j, k) of every node, node id order (z fastest)."""
    i, j, k = np.meshgrid(np.arange(m.nnx), np.arange(m.nny), np.arange(m.nnz), indexing="ij")
    return i.ravel(), j.ravel(), k.ravel()


def fill_isotropic(sig_real: np.ndarray, air: np.ndarray, omega0: float) -> np.ndarray:
    """Complex (npt, 6) tensor from per-node real isotropic sigma, air mask."""
    npt = sig_real.shape[0]
    s = np.zeros((npt, 6), dtype=np.complex128)
    im = f32(EPS0 * omega0)
    for c in (0, 3, 5):
        s[:, c] = np.where(air, 1j * im, sig_real + 1j * im)
    return s


def fill_tensor(sig6: np.ndarray, air: np.ndarray, omega0: float) -> np.ndarray:
    """Complex (npt, 6) tensor from a real (npt, 6) tensor; air overrides."""
    s = sig6.astype(np.complex128).copy()
    im = f32(EPS0 * omega0)
    for c in (0, 3, 5):
        s[:, c] = s[:, c] + 1j * im
    s[air] = 0
    for c in (0, 3, 5):
        s[air, c] = 1j * im
    return s


def mu_free(npt: int) -> np.ndarray:
    mu = np.zeros((npt, 6), dtype=np.float64)
    mu[:, 0] = mu[:, 3] = mu[:, 5] = 4.0 * PI * 1.0e-7
    return mu


def halfspace_block(nx=20, ny=20, nz_earth=16, nz_air=4, h=500.0, nodes=8, freqs=(1.0,),
                    sigma_bg=0.01, sigma_block=1.0, block=(4, 4, 2), block_top=1,
                    dirichlet=True, name="halfspace_block") -> MTModel:
    """Homogeneous half-space holding a conductive block (the 8-node Dirichlet
    workload: 20x20x16 cells of 500 m plus 4 air layers, 0.01 S/m with a 1 S/m
    2x2x1 km block whose top lies one cell below the surface)."""
    nord = 2 if nodes == 8 else 3
    nz = nz_earth + nz_air
    xp = _lattice(nx, h, nord)
    yp = _lattice(ny, h, nord)
    zl = _lattice(nz, h, nord)
    m = MTModel(nx, ny, nz, nodes, xp, yp, np.zeros(1), np.zeros((1, 6), complex), np.zeros((1, 6)),
                np.asarray(freqs, dtype=np.float64), dirichlet=dirichlet, name=name)
    i, j, k = node_grid(m)
    m.zp = zl[k].copy()
    ztop = nz_earth * h
    m.ztop = ztop
    air = m.zp > ztop + 1e-9
    sub = nord - 1
    bx0 = (nx - block[0]) // 2 * sub
    by0 = (ny - block[1]) // 2 * sub
    bz1 = (nz_earth - block_top) * sub
    bz0 = bz1 - block[2] * sub
    inblk = ((i >= bx0) & (i <= bx0 + block[0] * sub) & (j >= by0) & (j <= by0 + block[1] * sub)
             & (k >= bz0) & (k <= bz1))
    sr = np.where(inblk, sigma_block, sigma_bg)
    m.sigma = fill_isotropic(sr, air, m.omega(0))
    m.mu = mu_free(m.npt)
    m.nextd = 2
    m.nzl_top = min(2, nz_air)
    return m


def anisotropic_blocks(nx=64, ny=64, nz_earth=40, nz_air=8, h=500.0, nodes=20,
                       freqs=tuple(np.logspace(-3, 2, 8)), block=4, seed=2, full_tensor=False,
                       dirichlet=False, nextd=4, name="aniso_blocks") -> MTModel:
    """Block-wise anisotropic earth: each axis' log10 sigma ~ U[-3, 0] drawn per
    block x block x block cells (PCG64 seed); with full_tensor the principal
    values are rotated by ZYZ Euler angles drawn per block.  GPML on by default
    with nextd cells of stretching on +x, +y and the top."""
    nord = 2 if nodes == 8 else 3
    sub = nord - 1
    nz = nz_earth + nz_air
    xp = _lattice(nx, h, nord)
    yp = _lattice(ny, h, nord)
    zl = _lattice(nz, h, nord)
    m = MTModel(nx, ny, nz, nodes, xp, yp, np.zeros(1), np.zeros((1, 6), complex), np.zeros((1, 6)),
                np.asarray(freqs, dtype=np.float64), dirichlet=dirichlet, name=name)
    i, j, k = node_grid(m)
    m.zp = zl[k].copy()
    m.ztop = nz_earth * h
    air = m.zp > m.ztop + 1e-9
    rng = np.random.Generator(np.random.PCG64(seed))
    nbx, nby, nbz = (math.ceil(nx / block), math.ceil(ny / block), math.ceil(nz_earth / block))
    logs = rng.uniform(-3.0, 0.0, size=(nbx, nby, nbz, 3))
    ci = np.minimum(i // sub, nx - 1) // block
    cj = np.minimum(j // sub, ny - 1) // block
    ck = np.minimum(np.minimum(k // sub, nz - 1), nz_earth - 1) // block
    principal = 10.0 ** logs[ci, cj, ck]
    sig6 = np.zeros((m.npt, 6))
    if not full_tensor:
        sig6[:, 0], sig6[:, 3], sig6[:, 5] = principal[:, 0], principal[:, 1], principal[:, 2]
    else:
        ang = rng.uniform(0.0, 1.0, size=(nbx, nby, nbz, 3))
        al = 2 * PI * ang[..., 0]
        ga = 2 * PI * ang[..., 2]
        be = np.arccos(2 * ang[..., 1] - 1)
        R = _zyz(al, be, ga)[ci, cj, ck]
        D = np.zeros((m.npt, 3, 3))
        D[:, 0, 0], D[:, 1, 1], D[:, 2, 2] = principal[:, 0], principal[:, 1], principal[:, 2]
        S = R @ D @ np.transpose(R, (0, 2, 1))
        sig6[:, 0], sig6[:, 1], sig6[:, 2] = S[:, 0, 0], S[:, 0, 1], S[:, 0, 2]
        sig6[:, 3], sig6[:, 4], sig6[:, 5] = S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]
    m.sigma = fill_tensor(sig6, air, m.omega(0))
    m.mu = mu_free(m.npt)
    m.nextd = nextd
    m.nzl_top = nextd
    m.gpml_sch = 0
    m.a0, m.b0, m.nn = 1.0, 1.0, 2.0
    return m


def _zyz(al, be, ga):
    ca, sa, cb, sb, cg, sg = np.cos(al), np.sin(al), np.cos(be), np.sin(be), np.cos(ga), np.sin(ga)
    R = np.empty(al.shape + (3, 3))
    R[..., 0, 0] = ca * cb * cg - sa * sg
    R[..., 0, 1] = -ca * cb * sg - sa * cg
    R[..., 0, 2] = ca * sb
    R[..., 1, 0] = sa * cb * cg + ca * sg
    R[..., 1, 1] = -sa * cb * sg + ca * cg
    R[..., 1, 2] = sa * sb
    R[..., 2, 0] = -sb * cg
    R[..., 2, 1] = sb * sg
    R[..., 2, 2] = cb
    return R


def random_small(nodes=8, nx=3, ny=2, nz=2, seed=0, dirichlet=True, anisotropic=False,
                 full_tensor=False, topo=0.0, nextd=1, nz_air=1, freqs=(1.0, 10.0),
                 stretch=True, mu_contrast=False) -> MTModel:
    """Tiny randomized model for parity tests: non-uniform spacings, optional
    topography (z of each node column shifted smoothly), random conductivities."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nord = 2 if nodes == 8 else 3
    sub = nord - 1

    def axis(n):
        hs = rng.uniform(300.0, 900.0, size=n) if stretch else np.full(n, 500.0)
        pts = [0.0]
        for hh in hs:
            for s in range(1, sub + 1):
                pts.append(pts[-1] + hh / sub if s == 1 and sub == 2 else pts[-1] + hh / sub)
        return np.asarray(pts)

    xp, yp, zl = axis(nx), axis(ny), axis(nz)
    m = MTModel(nx, ny, nz, nodes, xp, yp, np.zeros(1), np.zeros((1, 6), complex), np.zeros((1, 6)),
                np.asarray(freqs, dtype=np.float64), dirichlet=dirichlet, name=f"small{nodes}")
    i, j, k = node_grid(m)
    z = zl[k].copy()
    if topo:
        # shift whole columns (node z depends on (i,j)); keeps elements right-handed
        zmax = zl[-1]
        bump = topo * np.sin(0.7 * xp[i] / 1000.0 + 0.3) * np.cos(0.5 * yp[j] / 1000.0)
        z = z + bump * (z / zmax)
    m.zp = z
    nz_earth = nz - nz_air
    m.ztop = zl[nz_earth * sub]
    air = (k > nz_earth * sub)
    sr = 10.0 ** rng.uniform(-3.0, 0.0, size=m.npt)
    if anisotropic or full_tensor:
        sig6 = np.zeros((m.npt, 6))
        p = 10.0 ** rng.uniform(-3.0, 0.0, size=(m.npt, 3))
        if full_tensor:
            al, be, ga = (rng.uniform(0, 2 * PI, m.npt), np.arccos(rng.uniform(-1, 1, m.npt)),
                          rng.uniform(0, 2 * PI, m.npt))
            R = _zyz(al, be, ga)
            D = np.zeros((m.npt, 3, 3))
            D[:, 0, 0], D[:, 1, 1], D[:, 2, 2] = p[:, 0], p[:, 1], p[:, 2]
            S = R @ D @ np.transpose(R, (0, 2, 1))
            sig6[:, 0], sig6[:, 1], sig6[:, 2] = S[:, 0, 0], S[:, 0, 1], S[:, 0, 2]
            sig6[:, 3], sig6[:, 4], sig6[:, 5] = S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]
        else:
            sig6[:, 0], sig6[:, 3], sig6[:, 5] = p[:, 0], p[:, 1], p[:, 2]
        m.sigma = fill_tensor(sig6, air, m.omega(0))
    else:
        m.sigma = fill_isotropic(sr, air, m.omega(0))
    m.mu = mu_free(m.npt)
    if mu_contrast:
        scale = rng.uniform(1.0, 3.0, size=m.npt)
        m.mu[:, 0] *= scale
        m.mu[:, 3] *= scale
        m.mu[:, 5] *= scale
    m.nextd = nextd
    m.nzl_top = max(1, min(nz_air, nextd))
    return m


def hills_tensor(nx=128, ny=128, nz_earth=88, nz_air=8, h=500.0, nodes=27, freqs=tuple(np.logspace(-3, 3, 16)),
                 n_hills=8, block=4, seed=5, dirichlet=False, nextd=4, name="hills_tensor") -> MTModel:
    """Full-tensor earth under topography: the surface is the sum of n_hills
    Gaussian hills (amplitude ~ U(0, 200] m, width ~ U[1, 5] km, centres ~ U
    over the inner domain, seeded), every node column shifted by the surface
    height times its relative height above the grid bottom (elements stay
    right-handed); sigma as in anisotropic_blocks(full_tensor=True); GPML on +x,
    +y and the top by default."""
    m = anisotropic_blocks(nx=nx, ny=ny, nz_earth=nz_earth, nz_air=nz_air, h=h, nodes=nodes, freqs=freqs,
                           block=block, seed=seed, full_tensor=True, dirichlet=dirichlet, nextd=nextd, name=name)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    amp = rng.uniform(0.0, 200.0, n_hills)
    amp = np.where(amp == 0.0, 200.0, amp)
    width = rng.uniform(1000.0, 5000.0, n_hills)
    lo, hi = nextd * h, (nx - nextd) * h
    cx = rng.uniform(lo, hi, n_hills)
    cy = rng.uniform(nextd * h, (ny - nextd) * h, n_hills)
    i, j, k = node_grid(m)
    x, y = m.xp[i], m.yp[j]
    surf = np.zeros(m.npt)
    for a, w, x0, y0 in zip(amp, width, cx, cy):
        surf += a * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2.0 * w * w))
    zmax = m.zp.max()
    m.zp = m.zp + surf * (m.zp / zmax)
    return m


def many_models(count=32, lo=(32, 32, 24), hi=(80, 80, 60), nz_air=4, h=500.0, freqs=tuple(np.logspace(-3, 2, 4)),
                seed=3, dirichlet=True) -> list:
    """The many-model workload: `count` models whose dimensions are drawn per
    model as lo + round(s * (hi - lo)) with s ~ U[0, 1] (PCG64 seed), element
    types alternating 8 / 20 nodes, full-tensor sigma per 4^3 cell block
    (principal values log-uniform in [1e-3, 1] S/m, ZYZ Euler angles), nz_air
    air layers on top of the earth, Dirichlet boundaries (or GPML)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for q in range(count):
        s = rng.uniform(0.0, 1.0)
        nx, ny, nz = (int(a + round(s * (b - a))) for a, b in zip(lo, hi))
        out.append(anisotropic_blocks(nx=nx, ny=ny, nz_earth=nz - nz_air, nz_air=nz_air, h=h,
                                      nodes=8 if q % 2 == 0 else 20, freqs=freqs, block=4,
                                      seed=int(rng.integers(1 << 31)), full_tensor=True, dirichlet=dirichlet,
                                      nextd=2 if dirichlet else 4, name=f"many_{q}"))
    return out
