"""Synthetic workload generators (host only): the many-model batch and the
27-node topography model keep the sizes and element types the measurements
and the full-size GPU tests state."""
import numpy as np

from movfem_amd import model as M


def edge_count(nx, ny, nz):
    return nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz


def test_many_models_sizes_and_types():
    ms = M.many_models(count=8, hi=(40, 40, 30))
    assert [m.nodes_per_elem for m in ms] == [8, 20] * 4
    for m in ms:
        assert 32 <= m.nx <= 40 and 32 <= m.ny <= 40 and 24 <= m.nz <= 30
        assert m.nx == m.ny and m.dirichlet and len(m.freqs) == 4
        s = m.sigma.reshape(m.npt, 6)
        earth = s[:, 0].real > 0
        # full tensor: off-diagonal components present in the earth, symmetric positive definite
        assert np.abs(s[earth][:, 1].real).max() > 0
        S = np.zeros((earth.sum(), 3, 3))
        S[:, 0, 0], S[:, 0, 1], S[:, 0, 2] = s[earth][:, 0].real, s[earth][:, 1].real, s[earth][:, 2].real
        S[:, 1, 1], S[:, 1, 2], S[:, 2, 2] = s[earth][:, 3].real, s[earth][:, 4].real, s[earth][:, 5].real
        S = S + np.triu(S, 1).transpose(0, 2, 1)
        ev = np.linalg.eigvalsh(S[:: max(1, len(S) // 2000)])
        assert ev.min() > 0 and ev.min() >= 0.99e-3 and ev.max() <= 1.01
    # same seed, same batch
    assert [(m.nx, m.nz) for m in M.many_models(count=8, hi=(40, 40, 30))] == [(m.nx, m.nz) for m in ms]


def test_hills_model_matches_stated_size():
    m = M.hills_tensor(nx=16, ny=16, nz_earth=10, nz_air=2)
    assert m.nodes_per_elem == 27 and m.nnx == 33 and m.npt == 33 * 33 * 25
    # the 27-node DOF count is the edge count of the twice-refined lattice
    assert edge_count(2 * 16, 2 * 16, 2 * 12) == 2 * 16 * 33 * 25 + 33 * 2 * 16 * 25 + 33 * 33 * 2 * 12
    full = (128, 128, 96)
    assert edge_count(*(2 * d for d in full)) == 38077120
    # topography: the surface is not flat, elements stay right-handed (z increasing along every column)
    z = m.zp.reshape(m.nnx, m.nny, m.nnz)
    assert np.ptp(z[:, :, -1]) > 0 and (np.diff(z, axis=2) > 0).all()
