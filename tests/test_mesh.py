"""Input readers and lattice generation (host code of libmovfem_gpu.so) against
the reference's own input() and grid_3d.

The golden vectors (tests/golden/mesh_*.npz, written by make_mesh_golden.py
from the reference build) hold what the reference's reader returned for each
3DEM_FEM.INP and the lattice its grid_3d built: g_xp, g_yp, g_nzl and the
shifted sites as arrays, g_zp, g_sigma and g_mu as SHA-256 digests of their
bytes.  The bar is bit-exact.  Cases: the reference's example input (flat
surface, 8-node elements) and the synthetic cases of tests/inputfiles.py.
One more synthetic case, absent from the fixtures, is run through the
reference build live when it is present.  PARAM.INP reads are checked against
values fixed by the file text.
"""
from __future__ import annotations

import hashlib
import pathlib

import numpy as np
import pytest

from movfem_amd.mesh import MeshInput, grid_3d, read_inputs, receivers, to_model
from movfem_amd.vfem import MvfError
from tests.inputfiles import CASES, case_files, make_case, render_3dem, render_param

GOLDEN = pathlib.Path(__file__).parent / "golden"
NAMES = ["sample", *CASES]


def _read_case(name, tmp_path):
    if name == "sample":
        return read_inputs(GOLDEN / "sample_3DEM_FEM.INP", GOLDEN / "sample_PARAM.INP")[0], [0.1]
    text, freqs = case_files(name)
    p = tmp_path / "3DEM_FEM.INP"
    p.write_text(text)
    return read_inputs(p)[0], freqs


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _same(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def _check_digest(z, key, ours):
    if _sha(ours) == str(z[key + "_sha256"]):
        return
    s = z[key + "_sample"]
    flat = np.ascontiguousarray(ours).ravel()
    step = max(1, len(flat) // 199)
    mine = flat[::step]
    n = min(len(s), len(mine))
    bad = np.flatnonzero(mine[:n] != s[:n])
    raise AssertionError(f"{key}: digest differs; sampled entries differing: {len(bad)} "
                         f"(first at {bad[:5] * step}: ours {mine[bad[:3]]} ref {s[bad[:3]]})")


@pytest.mark.parametrize("name", NAMES)
def test_reader_matches_reference(name, tmp_path):
    z = np.load(GOLDEN / f"mesh_{name}.npz")
    mi, _ = _read_case(name, tmp_path)
    nsr, nsf, nmax, imu, isig, mx, my, mz, mn = (int(v) for v in z["in_isz"])
    assert (len(mi.xsr), mi.nsf, mi.xto.shape[1], 0 if mi.ijmu is None else len(mi.ijmu), len(mi.ijsigma),
            len(mi.xm), len(mi.ym), mi.mz, mi.mn) == (nsr, nsf, nmax, imu, isig, mx, my, mz, mn)
    assert _same(np.array([mi.dx, mi.dy, mi.dz, mi.high, mi.depth, mi.sigma_bg]), z["in_rsz"])
    for k in ("xsr", "ysr", "zsr", "xm", "ym"):
        assert _same(getattr(mi, k), z["in_" + k]), k
    assert _same(mi.nsp, z["in_nsp"])
    assert _same(np.asfortranarray(mi.ijsigma).ravel(order="F"), z["in_ijsigma"])
    if imu:
        assert _same(np.asfortranarray(mi.ijmu).ravel(order="F"), z["in_ijmu"])
    # (nsf, nmax) and (ncomp, n) arrays in the reference's column-major order
    for k in ("xto", "yto", "zto"):
        _check_digest(z, "in_" + k, np.asfortranarray(getattr(mi, k)).ravel(order="F"))
    _check_digest(z, "in_zm", mi.zm)
    _check_digest(z, "in_sigma", mi.sigma)
    n = mx * my * mz
    _check_digest(z, "in_mu", mi.mu if imu else np.ones(n))  # imu = 0: in_mu(1, n) = 1


@pytest.mark.parametrize("name", NAMES)
def test_grid_matches_reference(name, tmp_path):
    z = np.load(GOLDEN / f"mesh_{name}.npz")
    mi, freqs = _read_case(name, tmp_path)
    g = grid_3d(mi, freqs)
    gnx, gny, gnz, nnx, nny, nnz, npt, nextd, nsf, nord, nsr = (int(v) for v in z["g_isz"])
    assert (g.nx + 1, g.ny + 1, g.nz + 1, len(g.xp), len(g.yp), len(g.zp) // (len(g.xp) * len(g.yp)), len(g.zp),
            g.nextd, len(g.nzl), len(g.xsr)) == (gnx, gny, gnz, nnx, nny, nnz, npt, nextd, nsf, nsr)
    assert nord == (2 if g.mn == 8 else 3)
    assert _same(np.array([g.ztop, g.x0, g.y0, g.z0, g.omega]), z["g_rsz"])
    for k in ("xp", "yp", "xsr", "ysr", "zsr"):
        assert _same(getattr(g, k), z["g_" + k]), k
    assert _same(g.nzl, z["g_nzl"])
    _check_digest(z, "g_zp", g.zp)
    _check_digest(z, "g_sigma", g.sigma)
    _check_digest(z, "g_mu", g.mu)


def test_grid_live_reference(tmp_path):
    """A case outside the fixtures (20-node elements, three interfaces of 40-75
    points, one conductivity component per axis) through the reference build."""
    refrun = pytest.importorskip("oracle.refrun")
    if not (pathlib.Path(refrun.__file__).parent / "_ref" / "libmovfem_ref.so").exists():
        pytest.skip("reference build absent (make -C oracle ref)")
    c = make_case(seed=205, mn=20, ninterf=3, npts=(75, 52, 40), isigma=3, sigma_bg=0.05, freqs=(0.7, 70.0),
                  dx=1500.0, dy=1100.0, dz=900.0)
    text = render_3dem(c, "tab")
    p = tmp_path / "3DEM_FEM.INP"
    p.write_text(text)
    mi, _ = read_inputs(p)
    g = grid_3d(mi, c["freqs"])
    r = refrun.run_reference_grid(text, c["freqs"])
    for k in ("xp", "yp", "zp", "sigma", "mu", "nzl", "xsr", "ysr", "zsr"):
        assert _same(getattr(g, k), r["g_" + k]), k
    assert _same(np.array([g.ztop, g.x0, g.y0, g.z0, g.omega]), r["g_rsz"])


def test_model_and_receivers(tmp_path):
    """The module state initialize_vfem hands on: GPML constants only for
    boundary 0, nzl of the top two slots locating the earth's surface."""
    mi, _ = read_inputs(GOLDEN / "sample_3DEM_FEM.INP")
    _, p = read_inputs(None, GOLDEN / "sample_PARAM.INP")
    g = grid_3d(mi, p.freqs)
    m = to_model(g, p)
    assert (m.nx, m.ny, m.nz, m.nodes_per_elem) == (58, 58, 43, 8)
    assert not m.dirichlet and (m.gpml_sch, m.a0, m.b0, m.nn) == (1, 1.0, 1.0, 2.0)
    assert m.nextd == 4 and m.nzl_top == 4 and m.ztop == 75870.0
    rx = receivers(g)
    assert rx["ke_top"] == 43 - 4 - 10 and len(rx["x"]) == 42
    # the surface layer of nodes sits at g_ztop (flat topography; the layer
    # increments are summed from the bottom, so to rounding)
    k = (rx["ke_top"]) * (m.nord - 1)
    assert np.abs(m.zp.reshape(m.nnx, m.nny, m.nnz)[:, :, k] - m.ztop).max() < 1e-9 * m.ztop
    # air above it: sigma holds only i*float32(omega*eps) on the diagonal
    s = m.sigma.reshape(m.nnx, m.nny, m.nnz, 6)
    ie = float(np.float32(8.854187817e-12 * 2 * np.pi * 0.1))
    assert np.all(s[:, :, k + 1:, 0] == 1j * ie) and np.all(s[:, :, :k + 1, 0].real == 0.01)


def test_param_sample():
    _, p = read_inputs(None, GOLDEN / "sample_PARAM.INP")
    assert list(p.freqs) == [0.1] and (p.boundary, p.inimod, p.h_sigma) == (0, 1, 0.01)
    assert list(p.l_sigma) == [0.01] and len(p.l_dz) == 0
    assert (p.gpml_sch, p.a0, p.b0, p.nn) == (1, 1.0, 1.0, 2.0)


def test_param_list_directed(tmp_path):
    """A READ continues over records until its list is filled and leaves the
    rest of its last record unread; an empty list consumes one record; the
    GPML lines are not read for a Dirichlet boundary."""
    text = ("freqs\n3\n1.0 2.5D-1\n1e3 ignored-rest\nbc\n1\nmodel\n3\nhalf\n2*0.02\nnl\n3\nsig\n0.01\n"
            "0.002, 0.3\ndz\n0.4 1.1 9.9\n")
    p = tmp_path / "PARAM.INP"
    p.write_text(text)
    _, r = read_inputs(None, p)
    assert list(r.freqs) == [1.0, 0.25, 1000.0]
    assert (r.boundary, r.inimod, r.h_sigma) == (1, 3, 0.02)
    assert list(r.l_sigma) == [0.01, 0.002, 0.3] and list(r.l_dz) == [0.4, 1.1]
    assert (r.gpml_sch, r.a0, r.b0, r.nn) == (0, 0.0, 0.0, 0.0)
    p.write_text(render_param([0.5, 5.0], boundary=0, inimod=2, l_sigma=(0.1,), gpml_sch=0, a0=2.0, b0=3.0, nn=1.5))
    _, r = read_inputs(None, p)
    assert list(r.freqs) == [0.5, 5.0] and len(r.l_dz) == 0 and (r.gpml_sch, r.a0, r.b0, r.nn) == (0, 2.0, 3.0, 1.5)


@pytest.mark.parametrize("text,what", [
    ("freqs\nx\n", "expected an integer"),
    ("freqs\n2\n1.0\n", "end of file"),
    ("freqs\n2\n1.0,,2.0\n", "null values"),
    ("freqs\n1\n1.0e\n", "expected a real"),
])
def test_param_errors(tmp_path, text, what):
    p = tmp_path / "PARAM.INP"
    p.write_text(text)
    with pytest.raises(MvfError, match=what):
        read_inputs(None, p)


def test_missing_file(tmp_path):
    with pytest.raises(MvfError, match="cannot open"):
        read_inputs(tmp_path / "nope.INP")


def test_grid_rejects_undefined_fits(tmp_path):
    """An interface of 12 points: the reference's qshep2 returns ier = 1 there
    (fewer points than the 13 least-squares equations need) and the reference
    goes on with undefined coefficients; the library refuses."""
    c = make_case(seed=7, mn=8, ninterf=1, npts=(12,))
    mi = MeshInput.build(c["dx"], c["dy"], c["dz"], c["high"], c["depth"], c["sigma_bg"], 8, c["sites"],
                         c["interfaces"], c["xm"], c["ym"], c["zm"], c["sigma"][..., 0], freqs=c["freqs"])
    with pytest.raises(MvfError, match="qshep2 ier=1"):
        grid_3d(mi)
    mi.mn = 9
    with pytest.raises(MvfError, match="mn must be"):
        grid_3d(mi)


def test_build_matches_file_path(tmp_path):
    """MeshInput.build (in-memory arrays) and the file reader give one lattice."""
    c = make_case(**CASES["topo8"][0])
    p = tmp_path / "3DEM_FEM.INP"
    p.write_text(render_3dem(c))
    mi_f, _ = read_inputs(p)
    mi_b = MeshInput.build(c["dx"], c["dy"], c["dz"], c["high"], c["depth"], c["sigma_bg"], c["mn"], c["sites"],
                           c["interfaces"], c["xm"], c["ym"], c["zm"], c["sigma"], ijsigma=c["ijsigma"])
    a, b = grid_3d(mi_f, c["freqs"]), grid_3d(mi_b, c["freqs"])
    for k in ("xp", "yp", "zp", "sigma", "mu", "nzl"):
        assert _same(getattr(a, k), getattr(b, k)), k
