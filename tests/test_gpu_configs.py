"""The benchmark configurations at full size on the device.

Config 1 (8-node, 20x20x20 cells of 500 m, half-space with a 1 S/m block,
1 Hz, Dirichlet): the device solve (direct-preconditioned GMRES on the
complex64 matrix the reference hands MUMPS) against SciPy SuperLU on the same
matrix (the MUMPS stand-in), read out by the oracle at 25 receivers on a 5x5
grid of 1 km spacing (rx_data / get_impedance / get_res_phase,
receiver_data.f90:111-133, solution.f90:456-491).  Bars: rho_a within 0.1 %,
phase within 0.05 degrees, secondary E on the gauge-determined DOFs within
1e-6 (relative L2).

Config 2 (20-node, 64x64x48 cells, GPML, diagonal anisotropic sigma) at the
lowest frequency, 1e-3 Hz:
  * DOF numbering bit-exact against the oracle, nne against the count formula
    2 E_grid + 2 F_grid, the CSR pattern equal to the coupling pattern derived
    on the host from the numbering (nnze included);
  * 1,100 sampled elements (GPML zones, the six faces, interior) with K_e, M_e
    and b_e within 1e-12 of the oracle's alocal/blocal;
  * the FP64 true relative residual of the device solve <= 1e-8 per
    polarization, recomputed on the host from the matrix global_vfem returns.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from movfem_amd import model as M
from movfem_amd.vfem import GpuVfem
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def config1_receivers(m):
    """5 x 5 receivers, 1 km apart, centred on the model."""
    cx, cy = 0.5 * (m.xp[0] + m.xp[-1]), 0.5 * (m.yp[0] + m.yp[-1])
    off = np.arange(-2, 3) * 1000.0
    xx, yy = np.meshgrid(cx + off, cy + off, indexing="ij")
    return xx.ravel(), yy.ravel()


def upper_matrix(ia, ja, a, nne):
    """CSR of what global_vfem returns: row-major full pattern, values on row <= col."""
    indptr = np.zeros(nne + 1, np.int64)
    np.cumsum(np.bincount(ia - 1, minlength=nne), out=indptr[1:])
    return sp.csr_matrix((a, ja - 1, indptr), shape=(nne, nne))


def sym_apply(U, x):
    """(U + U^T - diag U) x for the upper-triangle CSR U."""
    return U @ x + U.T @ x - U.diagonal() * x


def test_config1_receivers(need_gpu):
    from tests.gauge import determined_dofs

    m = M.halfspace_block()
    print("config 1: device solve vs SuperLU at 25 receivers", flush=True)
    g = GpuVfem(m)
    g.set_frequency(0)
    ia, ja, a, b = g.global_vfem()
    ia, ja, a = g.remove_zeros(ia, ja, a)
    n = g.nne
    U = sp.coo_matrix((a, (ia - 1, ja - 1)), shape=(n, n)).tocsr()
    A = (U + sp.triu(U, 1).T).tocsc()
    b = b.reshape(2, n)
    lu = spla.splu(A)
    xref = np.stack([lu.solve(b[d]) for d in range(2)])
    info = g.solve_frequencies([m.freqs[0]], method="gmres", precond="direct", tol=1e-8)
    assert info["converged"], info
    x = g.solution(2)
    for d in range(2):
        r = np.linalg.norm(b[d] - A @ x[d]) / np.linalg.norm(b[d])
        assert r <= 1e-8 * 1.0001, (d, r)
    mask = determined_dofs(m, g.gne())
    for d in range(2):
        e = np.linalg.norm(x[d][mask] - xref[d][mask]) / np.linalg.norm(xref[d][mask])
        print(f"pol {d}: secondary E rel L2 on determined DOFs {e:.2e}", flush=True)
        assert e <= 1e-6, (d, e)
    xs, ys = config1_receivers(m)
    ke_top = m.nz - 4
    out = g.readout(None, f=0, ke_top=ke_top, rx=xs, ry=ys)
    rho, phi = out[5], out[6]
    o = Oracle(m)
    o.set_frequency(0)
    oes, ohs = o.node_solution(xref.ravel())
    _, _, _, rho_ref, phi_ref = o.receivers(oes, ohs, ke_top, xs, ys)
    # the off-diagonal impedances carry the MT response; the diagonal ones
    # vanish by symmetry over the centred block, where the phase is noise
    for c in (1, 2):
        drho = np.abs(rho[:, c] - rho_ref[:, c]) / rho_ref[:, c]
        dphi = np.abs(phi[:, c] - phi_ref[:, c])
        print(f"Z component {c}: rho_a max rel diff {drho.max():.2e}, phase max diff {dphi.max():.2e} deg "
              f"(rho_a {rho_ref[:, c].min():.1f}..{rho_ref[:, c].max():.1f} ohm m)", flush=True)
        assert drho.max() <= 1e-3
        assert dphi.max() <= 0.05


def _flags(m, ie, je, ke):
    """GPML zone flags of element (ie, je, ke) (1-based): the reference's zones are
    the last nextd layers on +x, +y, +z (boundary_conds.f90:53-81), flag 1 = '+'."""
    if m.dirichlet:
        return 0, 0, 0
    return (int(ie > m.nx - m.nextd), int(je > m.ny - m.nextd), int(ke > m.nz - m.nextd))


def _ijk(m, e):
    return e // (m.ny * m.nz) + 1, (e // m.nz) % m.ny + 1, e % m.nz + 1


def sample_runs(m, rng, nruns=100, run=11):
    """Start elements of runs of `run` consecutive elements (reference loop order):
    40 % reaching into a GPML zone, 30 % on the six faces, 30 % interior."""
    starts = []
    nx, ny, nz, nd = m.nx, m.ny, m.nz, m.nextd
    while len(starts) < nruns:
        kind = len(starts) % 10
        ie, je = rng.integers(1, nx + 1), rng.integers(1, ny + 1)
        ke0 = rng.integers(1, nz - run + 2)
        if kind < 4:  # GPML: +x, +y or the top
            c = rng.integers(0, 3)
            if c == 0:
                ie = rng.integers(nx - nd + 1, nx + 1)
            elif c == 1:
                je = rng.integers(ny - nd + 1, ny + 1)
            else:
                ke0 = nz - run + 1
        elif kind < 7:  # faces
            f = rng.integers(0, 6)
            if f == 0:
                ie = 1
            elif f == 1:
                ie = nx
            elif f == 2:
                je = 1
            elif f == 3:
                je = ny
            elif f == 4:
                ke0 = 1
            else:
                ke0 = nz - run + 1
        else:
            ie = rng.integers(nd + 1, nx - nd + 1)
            je = rng.integers(nd + 1, ny - nd + 1)
        starts.append(((ie - 1) * ny + (je - 1)) * nz + (ke0 - 1))
    return sorted(set(int(s) for s in starts))


def test_config2_full_size(need_gpu):
    m = M.anisotropic_blocks(nx=64, ny=64, nz_earth=40, nz_air=8, h=500.0, nodes=20,
                             freqs=tuple(np.logspace(-3, 2, 8)), block=4, seed=2, nextd=4)
    print("config 2: setup", flush=True)
    g = GpuVfem(m)
    o = Oracle(m)
    # numbering and counts
    gne = g.gne()
    assert np.array_equal(gne, o.gne())
    nx, ny, nz = m.nx, m.ny, m.nz
    Eg = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    Fg = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
    assert g.nne == 2 * Eg + 2 * Fg == 2421088
    print("config 2: pattern", flush=True)
    rp, col, _, _ = g.operators()
    # coupling pattern from the numbering: every (row, col) sharing an element
    me = m.edges_per_elem
    rows = np.repeat(gne.astype(np.int64) - 1, me, axis=1).ravel()
    cols = np.tile(gne.astype(np.int64) - 1, (1, me)).ravel()
    key = np.unique(rows * g.nne + cols)
    del rows, cols
    assert key.size == g.nnze
    assert np.array_equal(np.diff(rp), np.bincount(key // g.nne, minlength=g.nne))
    assert np.array_equal(col, (key % g.nne).astype(np.int32))
    del key
    # sampled element matrices against the oracle (first frequency, first call:
    # element e is stretched with element e-1's zone flags)
    print("config 2: sampled elements", flush=True)
    g.set_frequency(0)
    o.set_frequency(0)
    w32 = float(np.float32(m.omega(0)))
    iu = np.triu_indices(me)
    rng = np.random.Generator(np.random.PCG64(22))
    starts = sample_runs(m, rng)
    nel, ea, eb, zones = 0, 0.0, 0.0, set()
    for s in starts:
        K, Mm, b = g.integrate(s, s + 11)
        o.set_in_pml(*(_flags(m, *_ijk(m, s - 1)) if s > 0 else (0, 0, 0)))
        for q in range(11):
            ie, je, ke = _ijk(m, s + q)
            zones.add(_flags(m, ie, je, ke))
            a_o, b_o = o.element(ie, je, ke)
            a_g = K[q] + 1j * (w32 * Mm[q])
            ea = max(ea, np.abs(a_g - a_o[iu]).max() / np.abs(a_o[iu]).max())
            eb = max(eb, np.abs(b[q] - b_o).max() / max(np.abs(b_o).max(), 1e-300))  # all-air elements: b = 0
            nel += 1
    print(f"config 2: {nel} elements, zones {sorted(zones)}, max rel err A {ea:.2e} b {eb:.2e}", flush=True)
    assert nel >= 1000 and len(zones) >= 4
    assert ea < 1e-12 and eb < 1e-12
    # the reference system of the first frequency and its device solve
    print("config 2: global_vfem + solve at 1e-3 Hz", flush=True)
    ia, ja, a, bb = g.global_vfem()
    x, info = g.solve(bb, method="gmres", precond="direct", tol=1e-8)
    print(f"config 2: {info['iterations']} iterations, library residuals {info['true_resid'][:2]}", flush=True)
    assert info["converged"], info
    U = upper_matrix(ia, ja, a, g.nne)
    del ia, ja, a
    n = g.nne
    for d in range(2):
        bd, xd = bb[d * n:(d + 1) * n], x[d * n:(d + 1) * n]
        r = np.linalg.norm(bd - sym_apply(U, xd)) / np.linalg.norm(bd)
        print(f"config 2: pol {d} host-recomputed true residual {r:.3e}", flush=True)
        assert r <= 1e-8 * 1.0001, (d, r)
