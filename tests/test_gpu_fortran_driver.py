"""The Fortran driver (ISO_C_BINDING, movfem_amd/fortran) runs the reference's
frequency loop shape on the GPU: global_vfem, zero removal and solve per
frequency, FP64 true residual <= 1e-8 on both polarizations.  The first
frequency goes through mvf_global_vfem + mvf_remove_zeros, the second through
the fused mvf_global_vfem_nz: both leave the same number of nonzero entries."""
import re
import subprocess

import pytest

from movfem_amd import build as B

pytestmark = pytest.mark.gpu


def test_fortran_driver(need_gpu):
    if not B.FDRIVER.exists():
        pytest.skip("Fortran driver not built (amdflang missing at build time)")
    r = subprocess.run([str(B.FDRIVER), "1.0", "0.1"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("freq")]
    assert len(lines) == 2
    nnz = [int(re.findall(r"nnz=\s*(\d+)", ln)[0]) for ln in lines]
    assert nnz[0] == nnz[1] > 0
    for ln in lines:
        assert "nne=   21660" in ln
        res = [float(x) for x in re.findall(r"resid=\s*(\S+)\s+(\S+)", ln)[0]]
        assert max(res) <= 1e-8
