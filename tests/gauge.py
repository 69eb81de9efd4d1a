"""Gauge-determined DOFs of a model.

The reference keeps only Re(sigma) in the mass term (integration.f90:211-238),
and air nodes carry a purely imaginary sigma, so elements whose nodes are all
air add curl-curl only.  Gradients of nodal functions supported inside that
region form a null space of the system matrix: there, any solver returns an
arbitrary gauge.  Edges of elements with at least one conducting node (the
earth and the first air layer) are determined; solutions are compared there.
"""
import numpy as np


def conducting_elements(m) -> np.ndarray:
    o = m.nord - 1
    cond = (m.sigma[:, [0, 3, 5]].real > 0).any(axis=1).reshape(m.nnx, m.nny, m.nnz)
    out = np.zeros(m.ne, bool)
    e = 0
    for i in range(m.nx):
        for j in range(m.ny):
            for k in range(m.nz):
                out[e] = cond[i * o:i * o + o + 1, j * o:j * o + o + 1, k * o:k * o + o + 1].any()
                e += 1
    return out


def determined_dofs(m, gne) -> np.ndarray:
    nne = int(gne.max())
    mask = np.zeros(nne, bool)
    ce = conducting_elements(m)
    g = gne[ce]
    mask[g[g > 0] - 1] = True
    return mask


def inner_dofs(m, gne) -> np.ndarray:
    """DOFs touched only by elements outside the GPML zones.  The reference's
    GPML is one-sided: zones are the last nextd element layers on +x, +y and
    +z (boundary_conds.f90:53-81).  Inside those zones the real-part-only
    stretch leaves modes with mass and almost no stiffness, which a solution at
    a given residual determines only weakly; the fields a survey reads lie
    outside them."""
    nne = int(gne.max())
    ie = np.repeat(np.arange(1, m.nx + 1), m.ny * m.nz)
    je = np.tile(np.repeat(np.arange(1, m.ny + 1), m.nz), m.nx)
    ke = np.tile(np.arange(1, m.nz + 1), m.nx * m.ny)
    if m.dirichlet:
        pml = np.zeros(m.ne, bool)
    else:
        pml = (ie > m.nx - m.nextd) | (je > m.ny - m.nextd) | (ke > m.nz - m.nextd)
    bad = np.zeros(nne + 1, bool)
    bad[np.maximum(gne[pml].ravel(), 0)] = True
    ok = np.zeros(nne + 1, bool)
    ok[np.maximum(gne[~pml].ravel(), 0)] = True
    out = ok & ~bad
    return out[1:]

