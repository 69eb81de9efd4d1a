"""The library's text writers (mvf_write_solution / _receivers / _surface)
against the reference's own writers (write_solution, write_receiver_data,
write_surface_data, solution.f90:73-205, receiver_data.f90:25-109) run by the
reference build on the same nodal fields, and against golden files for runs
without the reference.  From the same nodal fields the files must be
character-identical: formatted columns (es20.5, f20.5, A15, i5) and
list-directed records, whose reals take the reference build's digit choice."""
import pathlib
import tempfile

import numpy as np
import pytest

from movfem_amd import vfem
from oracle.oracle import Oracle, ref_available
from tests.outfiles import compare_files, receiver_heights
from tests.test_oracle_readout import CASES, air_layers, receivers_for, solve_oracle

GOLD = pathlib.Path(__file__).parent / "golden" / "out_hex8_small.npz"


def write_ours(m, es, hs, xs, ys, nzl_a, nzl_b, td):
    """The three files from nodal fields es, hs (2, npt, 3), with the oracle's
    impedance and receivers."""
    o = Oracle(m)
    o.set_frequency(0)
    nz, nr, nph = o.node_impedance(es, hs)
    ke_top = m.nz - nzl_a - nzl_b
    e, h, z, rho, phi = o.receivers(es, hs, ke_top, xs, ys)
    vfem.write_solution(td / "SOLUTION.OUT", 1, es, hs, nz, nr, nph)
    vfem.write_receivers(td / "RECEIVER.OUT", 1, 1, xs, ys, receiver_heights(m, xs, ys, ke_top), e, h, z, rho, phi)
    vfem.write_surface(td / "SURFACE.OUT", 1, 1, vfem.surface_nodes(m, nzl_a, nzl_b), es, hs, nz, nr, nph)
    return {n: (td / n).read_text() for n in ("SOLUTION.OUT", "RECEIVER.OUT", "SURFACE.OUT")}


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("name", ["hex8_dirichlet", "hex20_gpml_aniso"])
def test_writers_match_reference(name):
    from oracle.refrun import run_reference

    m = CASES[name]()
    o = Oracle(m)
    o.set_frequency(0)
    x = solve_oracle(o)
    xs, ys = receivers_for(m)
    nza = air_layers(m)
    nzl_a, nzl_b = m.nzl_top, nza - m.nzl_top
    r = run_reference(m, readout_x=x, receivers={"xy": np.stack([xs, ys], 1), "nzl": [nzl_a, nzl_b]},
                      write_files=True)
    with tempfile.TemporaryDirectory() as td:
        ours = write_ours(m, r["esol"], r["hsol"], xs, ys, nzl_a, nzl_b, pathlib.Path(td))
    for n, text in ours.items():
        exact, total = compare_files(text, r["files"][n])
        assert exact == total, (n, exact, total)


def test_writers_match_golden():
    z = np.load(GOLD)
    m = CASES["hex8_dirichlet"]()
    assert np.array_equal(m.zp, z["zp"])
    xs, ys = receivers_for(m)
    with tempfile.TemporaryDirectory() as td:
        ours = write_ours(m, z["esol"], z["hsol"], xs, ys, int(z["nzl_a"]), int(z["nzl_b"]), pathlib.Path(td))
    for n, text in ours.items():
        ref = bytes(z[n.replace(".", "_")]).decode()
        exact, total = compare_files(text, ref)
        assert exact == total, (n, exact, total)


def test_list_directed_records():
    """Records wrap before 80 columns, reals read back exactly."""
    with tempfile.TemporaryDirectory() as td:
        p = pathlib.Path(td) / "S"
        xs = np.array([0.0, 0.1, 1000.0, 1e-3, 1.5e-7, 1e20, -2.5, 1.2345678901234568e17, 1e100,
                       3.141592653589793, 0.0999999999, 1e15, 123456789012345.0, -0.0, 5e-324] * 3)
        surf = (len(xs), 1, xs, np.zeros(1), np.zeros(len(xs)), np.zeros(len(xs), np.int64))
        z0 = np.zeros(6, complex)
        vfem.write_surface(p, 1, 1, surf, z0, z0, np.zeros(4, complex), np.zeros(4), np.zeros(4))
        lines = p.read_text().splitlines()
    toks = []
    for ln in lines[2:]:
        assert ln.startswith(" ") and len(ln) <= 79
        if ln.strip() == "0.":
            break
        toks += ln.split()
    assert [float(t.replace("E", "e")) for t in toks] == list(xs)
    assert toks[:10] == ["0.", ".1", "1000.", "1.E-03", "1.5E-07", "1.E+20", "-2.5", "1.2345678901234568E+17",
                         "1.E+100", "3.141592653589793"]
    assert toks[10:15] == [".0999999999", "1.E+15", "123456789012345.", "-0.", "5.E-324"]


def test_list_directed_reals_match_compiler():
    """5,494 doubles (tests/golden/ld_reals.npz: 1e-300..1e300, short decimals,
    17-digit values, powers of two and ten with their neighbours, subnormals)
    print exactly as the reference's compiler prints them list-directed."""
    z = np.load(pathlib.Path(__file__).parent / "golden" / "ld_reals.npz")
    xs, want = z["x"], [str(t) for t in z["text"]]
    with tempfile.TemporaryDirectory() as td:
        p = pathlib.Path(td) / "S"
        surf = (len(xs), 1, xs, np.zeros(1), np.zeros(len(xs)), np.zeros(len(xs), np.int64))
        z0 = np.zeros(6, complex)
        vfem.write_surface(p, 1, 1, surf, z0, z0, np.zeros(4, complex), np.zeros(4), np.zeros(4))
        lines = p.read_text().splitlines()
    toks = []
    for ln in lines[2:]:
        toks += ln.split()
        if len(toks) >= len(xs):
            break
    bad = [(i, a, b) for i, (a, b) in enumerate(zip(toks, want)) if a != b]
    assert not bad, (len(bad), bad[:5])


# ----------------------------------------------------------------- PROBLEM.OUT
PCASES = ["hex8_dirichlet_topo_mu", "hex8_gpml_aniso", "hex20_gpml_tensor_topo", "hex27_gpml_mu_topo"]


def _problem_text(m):
    o = Oracle(m)
    o.set_frequency(0)
    ep, hp, src = o.problem_fields()
    with tempfile.TemporaryDirectory() as td:
        p = pathlib.Path(td) / "PROBLEM.OUT"
        vfem.write_problem(p, 1, ep, hp, src)
        return (ep, hp, src), p.read_text()


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("name", PCASES)
def test_problem_matches_reference(name):
    """The oracle's write_problem_elem values equal the reference's (the
    reference prints every double in a form that reads back exactly), and the
    library's writer reproduces the reference's PROBLEM.OUT character for
    character from them.  Cases: 8/20/27-node elements, topography, tensor
    conductivity, permeability contrasts (non-zero curl term)."""
    from oracle.refrun import run_reference
    from tests.outfiles import parse_problem
    from tests.parity_cases import cases

    m = cases()[name]
    r = run_reference(m, write_problem=True)
    (ep, hp, src), text = _problem_text(m)
    blocks = parse_problem(r["problem_out"])
    assert len(blocks) == 2
    for d, (head, arr) in enumerate(blocks):
        assert arr.shape == (m.npt, 10)
        assert np.all(arr[:, 0] == 0)
        assert np.array_equal(arr[:, 1:4], ep[d]) and np.array_equal(arr[:, 4:7], hp[d])
        assert np.array_equal(arr[:, 7:10], src[d])
    assert text == r["problem_out"]


def test_problem_matches_golden():
    from tests.parity_cases import cases

    ref = bytes(np.load(pathlib.Path(__file__).parent / "golden" / "out_problem_hex8_mu.npz")["PROBLEM_OUT"]).decode()
    _, text = _problem_text(cases()["hex8_dirichlet_topo_mu"])
    assert text == ref


def test_problem_inner_selection():
    """inner = 1: the rows of the selected nodes, in selection order."""
    from tests.outfiles import parse_problem
    from tests.parity_cases import cases

    m = cases()["hex8_gpml_aniso"]
    (ep, hp, src), _ = _problem_text(m)
    sel = np.array([5, 1, m.npt], np.int32)
    with tempfile.TemporaryDirectory() as td:
        p = pathlib.Path(td) / "P"
        vfem.write_problem(p, 3, ep, hp, src, sel=sel)
        blocks = parse_problem(p.read_text())
    for d, (head, arr) in enumerate(blocks):
        assert head.split()[-3:] == ["3", "f=", "3"]
        assert np.array_equal(arr[:, 1:4], ep[d][sel - 1]) and np.array_equal(arr[:, 7:10], src[d][sel - 1])
