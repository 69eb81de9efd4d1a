"""From input files to output files on the GPU.

* The Fortran program run_case (the reference's main program with every step
  on the library) runs the reference's own example input: 3DEM_FEM.INP holds a
  uniform 0.01 S/m half-space, PARAM.INP asks for 0.1 Hz with a GPML boundary.
  The 1-D answer is closed form: rho_a = 1/sigma = 100 ohm-m, a 45 degree
  phase, Zxx = Zyy = 0.  With 8-node elements the readout's H is the curl of
  the top earth cell's E (solution.f90:304-438), constant over the cell: the
  field half a cell (h/2 = 995 m) below the surface, which scales |Z|^2 by
  exp(h/delta) (delta = 15.9 km): rho_a = 100 exp(h/delta) = 113.3 ohm-m.
  RECEIVER.OUT must show that, within 3 %, at the sites within 15 km of the
  model's centre, with phases between 45 and 50 degrees.
* The Python path (movfem_amd.mesh.load_case + GpuVfem) on a synthetic case
  with topography on two interfaces: the device system equals the oracle's
  global_vfem on the generated lattice with the bars of the hand-built
  lattices (ia/ja exact, values within 1.2e-7 relative plus 1e-13 of the
  largest, RHS within 1e-12).
"""
import pathlib
import re
import shutil
import subprocess

import numpy as np
import pytest

from movfem_amd import build as B
from movfem_amd.mesh import load_case, receivers
from movfem_amd.vfem import GpuVfem
from oracle.oracle import Oracle
from oracle.refrun import parse_receiver_out
from tests.outfiles import parse_problem
from tests.inputfiles import CASES, make_case, render_3dem, render_param

pytestmark = pytest.mark.gpu
GOLDEN = pathlib.Path(__file__).parent / "golden"


def test_run_case_halfspace(need_gpu, tmp_path):
    if not B.FRUNCASE.exists():
        pytest.skip("run_case not built (amdflang missing at build time)")
    case = tmp_path / "case"
    case.mkdir()
    shutil.copy(GOLDEN / "sample_3DEM_FEM.INP", case / "3DEM_FEM.INP")
    shutil.copy(GOLDEN / "sample_PARAM.INP", case / "PARAM.INP")
    print("run_case on the example input", flush=True)
    r = subprocess.run([str(B.FRUNCASE), str(case)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("freq ")]
    assert len(lines) == 1
    res = [float(x) for x in re.findall(r"resid=\s*(\S+)\s+(\S+)", lines[0])[0]]
    assert max(res) <= 1e-8
    for name in ("PROBLEM.OUT", "SOLUTION.OUT", "RECEIVER.OUT", "SURFACE.OUT"):
        assert (tmp_path / name).stat().st_size > 0, name
    # PROBLEM.OUT: the inner region (nextd = 4 cells off each side, earth layers
    # from the fifth up to the surface) for both polarizations
    blocks = parse_problem((tmp_path / "PROBLEM.OUT").read_text())
    assert len(blocks) == 2 and blocks[0][1].shape == (51 * 51 * 26, 10)
    out = parse_receiver_out(tmp_path / "RECEIVER.OUT", 42)
    rho, phi, z = out["rx_rho"], out["rx_phi"], out["rx_z"]
    # sites: 21 along x at y = 50 km, then 21 along y at x = 50 km, 5 km apart
    centre = np.r_[7:14, 28:35]
    delta = np.sqrt(2.0 / (4e-7 * np.pi * 0.01 * 2 * np.pi * 0.1))
    expect = 100.0 * np.exp(1990.0 / delta)
    off, ph = rho[centre][:, 1:3], phi[centre][:, 1:3]
    print("rho_a", off.min(), off.max(), "expected", expect, "phase", ph.min(), ph.max())
    assert np.all(np.abs(off / expect - 1.0) < 0.03)
    assert np.all((ph > 45.0) & (ph < 50.0))
    # the diagonal impedances vanish next to the off-diagonal ones
    assert np.abs(z[centre][:, [0, 3]]).max() < 0.02 * np.abs(z[centre][:, 1:3]).min()


def test_input_files_to_device_system(need_gpu, tmp_path):
    c = make_case(**CASES["topo8"][0])
    (tmp_path / "3DEM_FEM.INP").write_text(render_3dem(c))
    (tmp_path / "PARAM.INP").write_text(render_param(c["freqs"], boundary=0, gpml_sch=1))
    m, mesh, p = load_case(tmp_path)
    print("device system of", m.nx, m.ny, m.nz, m.nodes_per_elem, flush=True)
    g = GpuVfem(m)
    o = Oracle(m)
    for fi in range(len(m.freqs)):
        g.set_frequency(fi)
        o.set_frequency(fi)
        ia, ja, a, b = g.global_vfem()
        ia_o, ja_o, a_o, b_o = o.global_vfem()
        assert np.array_equal(ia, ia_o) and np.array_equal(ja, ja_o)
        tol = 1.2e-7 * np.abs(a_o) + 1e-13 * np.abs(a_o).max()
        assert np.all(np.abs(a - a_o) <= tol), np.abs(a - a_o).max()
        assert np.linalg.norm(b - b_o) <= 1e-12 * np.linalg.norm(b_o)
    rx = receivers(mesh)
    assert rx["ke_top"] == m.nz - int(mesh.nzl[-1]) - int(mesh.nzl[-2])
