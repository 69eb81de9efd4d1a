"""Readout restatement (oracle node_solution / receivers) against the
reference's own solution.f90 and receiver_data.f90 compiled by oracle/Makefile
(oracle/_ref), on solutions of the restated system (SciPy SuperLU).

node_solution: total E and H at every lattice node, first element in loop
order wins (solution.f90:18-49, 207-300, 304-438).  Receivers: bilinear
interpolation on the top face of the top earth element, Z = E H^-1, rho_a,
phase (receiver_data.f90:111-253, solution.f90:440-505); the reference prints
receivers with 5-6 significant digits (RECEIVER.OUT formats es20.5, f20.5),
which sets the receiver tolerances.  Also pinned by physics: on a homogeneous
half-space the apparent resistivity is 1/sigma and the phase 45 degrees."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from movfem_amd import model as M
from oracle.oracle import Oracle, ref_available

needs_ref = pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not available")


def solve_oracle(o):
    ia, ja, a, b = o.global_vfem()
    ia, ja, a = Oracle.remove_zeros(ia, ja, a)
    n = o.nne
    U = sp.coo_matrix((a, (ia - 1, ja - 1)), shape=(n, n)).tocsr()
    A = (U + sp.triu(U, 1).T).tocsc()
    lu = spla.splu(A)
    b = b.reshape(2, n)
    return np.concatenate([lu.solve(b[d]) for d in range(2)])


CASES = {
    "hex8_dirichlet": lambda: M.random_small(8, 6, 6, 4, seed=31, nz_air=2, freqs=(1.0,)),
    "hex8_gpml_aniso": lambda: M.anisotropic_blocks(nx=8, ny=8, nz_earth=4, nz_air=2, nodes=8, freqs=(3.0,),
                                                    block=2, seed=7, nextd=2),
    "hex20_gpml_aniso": lambda: M.anisotropic_blocks(nx=6, ny=6, nz_earth=3, nz_air=2, nodes=20, freqs=(0.5,),
                                                     block=2, seed=8, nextd=2),
    "hex27_dirichlet_mu": lambda: M.random_small(27, 4, 4, 3, seed=32, nz_air=2, freqs=(2.0,), mu_contrast=True),
}


def air_layers(m):
    """Element layers whose bottom lies at or above the earth surface."""
    o = m.nord - 1
    zc = m.zp[:m.nnz:o]  # first node column, corner lattice
    return int(np.sum(zc[:-1] >= m.ztop - 1e-9))


def receivers_for(m):
    """Receiver positions strictly inside the elements the reference searches
    (ie, je in [nextd+1, nx-nextd])."""
    o = m.nord - 1
    xs, ys = [], []
    for a in range(m.nextd + 1, m.nx - m.nextd + 1):
        for c in range(m.nextd + 1, m.ny - m.nextd + 1):
            x0, x1 = m.xp[(a - 1) * o], m.xp[a * o]
            y0, y1 = m.yp[(c - 1) * o], m.yp[c * o]
            xs.append(x0 + 0.3 * (x1 - x0))
            ys.append(y0 + 0.6 * (y1 - y0))
    return np.array(xs), np.array(ys)


@needs_ref
@pytest.mark.parametrize("name", sorted(CASES))
def test_readout_matches_reference(name):
    from oracle.refrun import run_reference

    m = CASES[name]()
    o = Oracle(m)
    o.set_frequency(0)
    x = solve_oracle(o)
    es, hs = o.node_solution(x)
    xs, ys = receivers_for(m)
    nza = air_layers(m)
    nzl_a, nzl_b = m.nzl_top, nza - m.nzl_top
    r = run_reference(m, vfem_freqs=[1], readout_x=x, receivers={"xy": np.stack([xs, ys], 1),
                                                                  "nzl": [nzl_a, nzl_b]})
    for got, ref in ((es, r["esol"]), (hs, r["hsol"])):
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-10 * scale
    ke_top = m.nz - nza
    e, h, z, rho, phi = o.receivers(es, hs, ke_top, xs, ys)
    assert np.allclose(e, r["rx_e"], rtol=2e-5, atol=1e-5 * np.abs(r["rx_e"]).max())
    assert np.allclose(h, r["rx_h"], rtol=2e-5, atol=1e-5 * np.abs(r["rx_h"]).max())
    assert np.allclose(z, r["rx_z"], rtol=2e-5, atol=1e-5 * np.abs(r["rx_z"]).max())
    assert np.allclose(rho, r["rx_rho"], rtol=1e-5, atol=2e-5)
    assert np.allclose(phi, r["rx_phi"], rtol=1e-5, atol=2e-5)


def edge_sites(m):
    """Sites on the inner cells' lower and upper edges and beyond them, after
    one inside."""
    q = m.nord - 1
    xa, xb = m.xp[m.nextd * q], m.xp[(m.nx - m.nextd) * q]
    ya, yb = m.yp[m.nextd * q], m.yp[(m.ny - m.nextd) * q]
    xm, ym = 0.5 * (xa + xb) + 1.0, 0.5 * (ya + yb) + 1.0
    return (np.array([xm, xa, xb, xb + 300.0, xm, xm, xa - 200.0, xm]),
            np.array([ym, ym, ym, ya, yb, ya - 50.0, yb + 10.0, ya]))


@needs_ref
def test_edge_receivers_match_reference():
    """Sites on and beyond the inner cells' edges: get_rx_element finds no
    cell for x (then y is not searched) or for y, and rx_data goes on with the
    previous receiver's indices, extrapolating from that element (the
    reference's example input has such sites at its model's edges)."""
    from oracle.refrun import run_reference

    m = CASES["hex8_dirichlet"]()
    o = Oracle(m)
    o.set_frequency(0)
    x = solve_oracle(o)
    es, hs = o.node_solution(x)
    xs, ys = edge_sites(m)
    nza = air_layers(m)
    nzl_a, nzl_b = m.nzl_top, nza - m.nzl_top
    r = run_reference(m, vfem_freqs=[1], readout_x=x, receivers={"xy": np.stack([xs, ys], 1), "nzl": [nzl_a, nzl_b]})
    e, h, z, rho, phi = o.receivers(es, hs, m.nz - nza, xs, ys)
    assert np.allclose(e, r["rx_e"], rtol=2e-5, atol=1e-5 * np.abs(r["rx_e"]).max())
    assert np.allclose(h, r["rx_h"], rtol=2e-5, atol=1e-5 * np.abs(r["rx_h"]).max())
    assert np.allclose(z, r["rx_z"], rtol=2e-5, atol=1e-5 * np.abs(r["rx_z"]).max())
    assert np.allclose(rho, r["rx_rho"], rtol=1e-5, atol=2e-5)
    # the first site outside every inner cell has nothing to inherit
    with pytest.raises(RuntimeError):
        o.receivers(es, hs, m.nz - nza, xs[[3, 0]], ys[[3, 0]])
