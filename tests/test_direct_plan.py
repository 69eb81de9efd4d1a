"""Symbolic analysis of the device sparse direct factorization, on the host.

The plan is built from the DOF numbering (the oracle's restatement of c_gne*,
global_assembly.f90:183-475) and the element coupling pattern.  The C++
builder itself checks that every coupled DOF of a front's owned DOFs is in
the front and that every update DOF reaches the parent front; here the plan
must also eliminate every DOF exactly once and its front sizes must follow
the nested-dissection geometry: the root separator of a box bisected along x
is the DOF set of one yz plane of the lattice."""
import ctypes as C

import numpy as np
import pytest

from movfem_amd import _lib as L
from movfem_amd import model as M
from oracle.oracle import Oracle


def pattern(gne, nne):
    """Full CSR pattern of the element couplings (0-based)."""
    rows, cols = [], []
    for g in gne:
        d = g[g > 0] - 1
        r, c = np.meshgrid(d, d, indexing="ij")
        rows.append(r.ravel())
        cols.append(c.ravel())
    key = np.unique(np.concatenate(rows).astype(np.int64) * nne + np.concatenate(cols))
    r, c = key // nne, key % nne
    rowptr = np.zeros(nne + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    return np.cumsum(rowptr), c.astype(np.int32)


def plan_stats(m, gne, nne):
    rowptr, col = pattern(gne, nne)
    st = np.zeros(10)
    g = np.ascontiguousarray(gne, np.int32)
    rc = L.lib().mvf_direct_plan_stats(m.nx, m.ny, m.nz, gne.shape[1], g.ctypes.data_as(C.POINTER(C.c_int32)), nne,
                                       rowptr.ctypes.data_as(C.POINTER(C.c_int64)),
                                       col.ctypes.data_as(C.POINTER(C.c_int32)),
                                       st.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, rc
    return dict(zip(["fronts", "depth", "max_n1", "max_n2", "max_m", "fac", "flops", "sum_n1", "sec", "upd"], st))


@pytest.mark.parametrize("nodes,dirichlet", [(8, True), (8, False), (20, False), (27, True)])
def test_plan_eliminates_every_dof_once(nodes, dirichlet):
    m = M.random_small(nodes, 8, 6, 5, seed=3, nz_air=1, dirichlet=dirichlet, nextd=1) if not dirichlet else \
        M.random_small(nodes, 8, 6, 5, seed=3, nz_air=1)
    o = Oracle(m)
    gne = o.gne()
    s = plan_stats(m, gne, o.nne)
    assert s["sum_n1"] == o.nne
    assert s["fronts"] >= 3 and s["max_m"] <= o.nne and s["flops"] > 0
    # root: bisection of the 8 x 6 x 5 box along x at ie = 4; its owned DOFs are
    # those shared by elements on both sides of that plane
    ie = np.repeat(np.arange(m.nx), m.ny * m.nz)
    left = set(gne[ie < 4].ravel()) - {0}
    right = set(gne[ie >= 4].ravel()) - {0}
    sep = {d for d in left & right if d > 0}
    assert s["max_n1"] >= len(sep)


def test_plan_front_sizes_on_a_cube():
    """8-node 8x8x8, Dirichlet: the root separator is the interior of the
    x = 4 plane: edges along y and z on it, minus Dirichlet boundary edges:
    y-edges 8 x 7 interior... counted directly from the numbering."""
    m = M.random_small(8, 8, 8, 8, seed=4, nz_air=1)
    o = Oracle(m)
    gne = o.gne()
    s = plan_stats(m, gne, o.nne)
    ie = np.repeat(np.arange(8), 64)
    sep = (set(gne[ie == 3].ravel()) & set(gne[ie == 4].ravel())) - {0}
    sep = {d for d in sep if d > 0}
    # 8-node plane x=4 with Dirichlet sides: y-edges (j in 0..7, k interior 1..7) and
    # z-edges (j interior 1..7, k in 0..7): 8*7 + 7*8 = 112
    assert len(sep) == 112
    assert s["max_n1"] == 112
