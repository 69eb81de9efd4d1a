"""Small models shared by the parity tests (seeded; the oracle finishes each in
well under a second)."""
from movfem_amd import model as M


def cases():
    return {
        "hex8_dirichlet_topo_mu": M.random_small(8, 3, 2, 3, seed=7, topo=300.0, mu_contrast=True, nz_air=1),
        "hex8_gpml_aniso": M.random_small(8, 3, 3, 3, seed=4, dirichlet=False, nextd=1, anisotropic=True),
        "hex20_dirichlet": M.random_small(20, 2, 2, 3, seed=2, nz_air=1),
        "hex20_gpml_tensor_topo": M.random_small(20, 3, 2, 3, seed=5, dirichlet=False, nextd=1, full_tensor=True,
                                                 topo=200.0),
        "hex27_dirichlet_tensor": M.random_small(27, 2, 2, 3, seed=3, full_tensor=True, nz_air=1),
        "hex27_gpml_mu_topo": M.random_small(27, 2, 3, 2, seed=6, dirichlet=False, nextd=1, full_tensor=True,
                                             topo=150.0, mu_contrast=True),
    }
