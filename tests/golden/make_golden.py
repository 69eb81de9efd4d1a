"""Writes golden vectors from the reference build (oracle/_ref, compiled from
the reference's own Fortran modules by oracle/Makefile).

Run in the container that holds the reference sources:
    make -C oracle ref && python -m tests.golden.make_golden
out_problem_hex8_mu.npz holds the reference's PROBLEM.OUT for one parity case.
Each model fixture stores the model inputs (lattice, sigma, mu, boundary settings)
and the reference's outputs: gne, nzindx, alocal/blocal of every element and
global_vfem (ia, ja, a, b) for the first two frequencies, plus g_sigma after
update_sigma for frequency 2.  The Dirichlet inimod 2/3 fixtures add the
boundary 1-D model (h_sigma, l_sigma, l_dz) and node_solution's E and H for a
seeded solution vector (readout_x, esol, hsol).

    python -m tests.golden.make_golden [name ...]   # default: every fixture
"""
from __future__ import annotations

import pathlib

import numpy as np

from movfem_amd import model as M
from oracle.refrun import run_reference

HERE = pathlib.Path(__file__).resolve().parent


def golden_cases():
    return {
        "g8_dirichlet_topo": M.random_small(8, 2, 2, 3, seed=31, topo=250.0, nz_air=1, freqs=(0.5, 7.0)),
        "g8_gpml_mu": M.random_small(8, 3, 2, 3, seed=32, dirichlet=False, nextd=1, mu_contrast=True,
                                     freqs=(2.0, 0.05)),
        "g20_gpml_tensor": M.random_small(20, 2, 2, 2, seed=33, dirichlet=False, nextd=1, full_tensor=True,
                                          topo=120.0, freqs=(1.0, 10.0)),
        "g27_dirichlet_aniso": M.random_small(27, 2, 2, 2, seed=34, anisotropic=True, nz_air=1, freqs=(3.0, 30.0)),
        "g8_dirichlet_halfspace_bd": _bd(M.random_small(8, 3, 2, 3, seed=35, topo=180.0, nz_air=1,
                                                        freqs=(0.2, 25.0)), 2, h_sigma=0.004),
        "g20_dirichlet_layered_bd": _bd(M.random_small(20, 2, 2, 3, seed=36, nz_air=1, freqs=(1.0, 0.01)), 3,
                                        l_sigma=(0.02, 0.0015, 0.4), l_dz=(0.35, 0.9)),
    }


def _bd(m, inimod, **kw):
    m.inimod = inimod
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def readout_vector(nne, seed=77):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(2 * nne) + 1j * rng.standard_normal(2 * nne)


def write_outfiles_golden():
    """out_hex8_small.npz: the reference's SOLUTION.OUT, RECEIVER.OUT and
    SURFACE.OUT for the 8-node Dirichlet readout case, with the nodal fields
    they were written from."""
    from oracle.oracle import Oracle
    from tests.test_oracle_readout import CASES, air_layers, receivers_for, solve_oracle

    m = CASES["hex8_dirichlet"]()
    o = Oracle(m)
    o.set_frequency(0)
    x = solve_oracle(o)
    xs, ys = receivers_for(m)
    nza = air_layers(m)
    nzl_a, nzl_b = m.nzl_top, nza - m.nzl_top
    r = run_reference(m, readout_x=x, receivers={"xy": np.stack([xs, ys], 1), "nzl": [nzl_a, nzl_b]},
                      write_files=True)
    files = {n.replace(".", "_"): np.frombuffer(r["files"][n].encode(), np.uint8) for n in r["files"]}
    np.savez_compressed(HERE / "out_hex8_small.npz", esol=r["esol"], hsol=r["hsol"], zp=m.zp, nzl_a=nzl_a,
                        nzl_b=nzl_b, **files)
    print("out_hex8_small", {k: len(v) for k, v in files.items()})


def write_problem_golden():
    """out_problem_hex8_mu.npz: the reference's PROBLEM.OUT (write_problem_ugrid,
    every node, both polarizations) for the 8-node Dirichlet case with topography
    and a permeability contrast (tests/parity_cases.py), at its first frequency."""
    from tests.parity_cases import cases

    r = run_reference(cases()["hex8_dirichlet_topo_mu"], write_problem=True)
    np.savez_compressed(HERE / "out_problem_hex8_mu.npz",
                        PROBLEM_OUT=np.frombuffer(r["problem_out"].encode(), np.uint8))
    print("out_problem_hex8_mu", len(r["problem_out"]))


def main(names=()):
    if not names or "out_hex8_small" in names:
        write_outfiles_golden()
    if not names or "out_problem_hex8_mu" in names:
        write_problem_golden()
    for name, m in golden_cases().items():
        if names and name not in names:
            continue
        els = [(i, j, k) for i in range(1, m.nx + 1) for j in range(1, m.ny + 1) for k in range(1, m.nz + 1)]
        bd = m.dirichlet and m.inimod in (2, 3)
        x = None
        if bd:
            x = readout_vector(int(run_reference(m)["nne"]))
        r = run_reference(m, elements=els, vfem_freqs=[1, 2], sigma_after=[], readout_x=x)
        # sigma after update_sigma at frequency 2 (separate process: state order)
        r2 = run_reference(m, sigma_after=[2])
        out = dict(nx=m.nx, ny=m.ny, nz=m.nz, mn=m.nodes_per_elem, xp=m.xp, yp=m.yp, zp=m.zp, sigma=m.sigma, mu=m.mu,
                   freqs=m.freqs, dirichlet=int(m.dirichlet), inimod=m.inimod, gpml_sch=m.gpml_sch, a0=m.a0,
                   b0=m.b0, nn=m.nn, nextd=m.nextd, nzl_top=m.nzl_top, ztop=m.ztop,
                   gne=r["gne"], nzindx=r["nzindx"], nne=r["nne"], nnze=r["nnze"], aloc=r["aloc"], bloc=r["bloc"],
                   sigma_2=r2["sigma_2"])
        if bd:
            out.update(h_sigma=m.h_sigma, l_sigma=np.asarray(m.l_sigma, np.float64),
                       l_dz=np.asarray(m.l_dz, np.float64), readout_x=x, esol=r["esol"], hsol=r["hsol"])
        for fi in (1, 2):
            for k in ("ia", "ja", "a", "b"):
                out[f"{k}_{fi}"] = r[f"{k}_{fi}"]
        np.savez_compressed(HERE / f"{name}.npz", **out)
        print(name, "nne", int(r["nne"]), "nnze", int(r["nnze"]))


def load_model(z) -> M.MTModel:
    m = M.MTModel(int(z["nx"]), int(z["ny"]), int(z["nz"]), int(z["mn"]), z["xp"], z["yp"], z["zp"], z["sigma"],
                  z["mu"], z["freqs"], dirichlet=bool(z["dirichlet"]), inimod=int(z["inimod"]),
                  gpml_sch=int(z["gpml_sch"]), a0=float(z["a0"]), b0=float(z["b0"]), nn=float(z["nn"]),
                  nextd=int(z["nextd"]), nzl_top=int(z["nzl_top"]), ztop=float(z["ztop"]))
    if "h_sigma" in z.files:
        m.h_sigma = float(z["h_sigma"])
        m.l_sigma = tuple(float(v) for v in z["l_sigma"])
        m.l_dz = tuple(float(v) for v in z["l_dz"])
    return m


if __name__ == "__main__":
    import sys

    main(tuple(sys.argv[1:]))
