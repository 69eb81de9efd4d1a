"""From input files to output files on the GPU.

* The Fortran program run_case (the reference's main program with every step
  on the library) runs the reference's own example input: 3DEM_FEM.INP holds a
  uniform 0.01 S/m half-space, PARAM.INP asks for 0.1 Hz with a GPML boundary.
  There the 1-D answer is closed form: rho_a = 1/(sigma) = 100 ohm-m and a
  45 degree phase for Zxy and Zyx, with Zxx = Zyy = 0.  RECEIVER.OUT must
  show that to the discretisation error (2000 m cells, skin depth 15.9 km).
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
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("freq")]
    assert len(lines) == 1
    res = [float(x) for x in re.findall(r"resid=\s*(\S+)\s+(\S+)", lines[0])[0]]
    assert max(res) <= 1e-8
    for name in ("SOLUTION.OUT", "RECEIVER.OUT", "SURFACE.OUT"):
        assert (tmp_path / name).stat().st_size > 0, name
    out = parse_receiver_out(tmp_path / "RECEIVER.OUT", 42)
    rho, phi, z = out["rx_rho"], out["rx_phi"], out["rx_z"]
    # off-diagonal apparent resistivity and phase of the half-space
    off = np.concatenate([rho[:, 1], rho[:, 2]])
    ph = np.concatenate([phi[:, 1], phi[:, 2]])
    print("rho_a", off.min(), np.median(off), off.max(), "phase", ph.min(), np.median(ph), ph.max())
    assert abs(np.median(off) - 100.0) < 3.0
    assert np.all(np.abs(off - 100.0) < 10.0)
    assert abs(np.median(ph) - 45.0) < 1.0
    assert np.all(np.abs(ph - 45.0) < 3.0)
    # the diagonal impedances vanish next to the off-diagonal ones
    zoff = np.abs(z[:, 1:3]).min()
    assert np.abs(z[:, [0, 3]]).max() < 0.05 * zoff


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
