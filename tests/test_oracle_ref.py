"""The CPU oracle against the reference's own modules, live (oracle/_ref).

Skipped where the reference build is absent (on a machine without the
reference sources); tests/test_oracle_golden.py carries the same checks on
committed vectors."""
import numpy as np
import pytest

from oracle.oracle import Oracle, ref_available
from tests.parity_cases import cases

pytestmark = pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")

CASES = cases()


def _rel(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s else 1.0)


@pytest.mark.parametrize("name", sorted(CASES))
def test_against_reference(name):
    from oracle.refrun import run_reference

    m = CASES[name]
    els = [(i, j, k) for i in range(1, m.nx + 1) for j in range(1, m.ny + 1) for k in range(1, m.nz + 1)]
    r = run_reference(m, elements=els, vfem_freqs=[1, 2])
    o = Oracle(m)
    assert np.array_equal(o.gne(), r["gne"])
    assert np.array_equal(o.nzindx(), r["nzindx"])
    o.set_frequency(0)
    for t, (ie, je, ke) in enumerate(els):
        a, b = o.element(ie, je, ke)
        assert _rel(a, r["aloc"][t]) < 1e-13
        assert _rel(b, r["bloc"][t]) < 1e-13
    for fi in (1, 2):
        o.set_frequency(fi - 1)
        ia, ja, a, b = o.global_vfem()
        assert np.array_equal(ia, r[f"ia_{fi}"]) and np.array_equal(ja, r[f"ja_{fi}"])
        ar = r[f"a_{fi}"]
        assert np.all(np.abs(a - ar) <= 1.2e-7 * np.abs(ar) + 1e-14 * np.abs(ar).max())
        assert _rel(b, r[f"b_{fi}"]) < 1e-13
