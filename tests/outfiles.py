"""Comparison of the reference's text output files (SOLUTION.OUT, RECEIVER.OUT,
SURFACE.OUT) written by two sources.

Lines must have the same length and tokens.  A token pair passes when the
strings are equal, or when both read as numbers and agree to rel_tol (0: the
same double).  Equal doubles give identical text (the library prints, among
the shortest decimals that read back, the middle one, as the reference build
does); the tolerances serve files written from device-computed values."""
import numpy as np


def _num(t):
    try:
        return float(t.replace("E", "e"))
    except ValueError:
        return None


def compare_files(ours: str, ref: str, rel_tol=0.0, abs_tol=0.0, phase_branch=False):
    """phase_branch: also accept values 180 apart (the phase is atan(Im Z / Re Z),
    solution.f90:486, which jumps between +-90 degrees where Re Z is at rounding
    level and its sign depends on the summation order)."""
    a, b = ours.splitlines(), ref.splitlines()
    assert len(a) == len(b), (len(a), len(b))
    exact = 0
    for i, (x, y) in enumerate(zip(a, b)):
        if x == y:
            exact += 1
            continue
        tx, ty = x.split(), y.split()
        assert len(tx) == len(ty), (i, x, y)
        # formatted records keep their field widths; list-directed ones (a single
        # leading blank) may differ in length with the last-digit choice
        assert len(x) == len(y) or (x.startswith(" ") and not x.startswith("  ")), (i, x, y)
        for p, q in zip(tx, ty):
            if p == q:
                continue
            u, v = _num(p), _num(q)
            assert u is not None and v is not None, (i, p, q)
            tol = max(rel_tol * max(abs(u), abs(v)), abs_tol)
            assert abs(u - v) <= tol or (phase_branch and abs(abs(u - v) - 180.0) <= tol), (i, p, q)
    return exact, len(a)


def receiver_heights(m, xs, ys, ke_top):
    """interpolate_rx's surface height (receiver_data.f90:211-225) in its own
    operation order."""
    o = m.nord - 1
    out = []
    for x, y in zip(xs, ys):
        ie = max(a for a in range(m.nextd + 1, m.nx - m.nextd + 1) if m.xp[(a - 1) * o] <= x)
        je = max(b for b in range(m.nextd + 1, m.ny - m.nextd + 1) if m.yp[(b - 1) * o] <= y)
        x1, x2, y1, y2 = m.xp[(ie - 1) * o], m.xp[ie * o], m.yp[(je - 1) * o], m.yp[je * o]
        k = ke_top * o

        def zid(i, j):
            return m.zp[(i * m.nny + j) * m.nnz + k]

        z11, z12 = zid((ie - 1) * o, (je - 1) * o), zid((ie - 1) * o, je * o)
        z21, z22 = zid(ie * o, (je - 1) * o), zid(ie * o, je * o)
        zx1 = ((x2 - x) / (x2 - x1)) * z11 + ((x - x1) / (x2 - x1)) * z21
        zx2 = ((x2 - x) / (x2 - x1)) * z12 + ((x - x1) / (x2 - x1)) * z22
        out.append(((y2 - y) / (y2 - y1)) * zx1 + ((y - y1) / (y2 - y1)) * zx2)
    return np.array(out)


def parse_problem(text):
    """PROBLEM.OUT blocks (write_problem_ugrid, problem.f90:188-240): a list of
    (header line, complex array (nrow, 10)): psigma, ep(3), hp(3), src(3) per
    row, records joined over their line breaks."""
    import re

    blocks, lines, i = [], text.split("\n"), 0
    while i < len(lines):
        if not lines[i].strip().startswith("problem main functions"):
            i += 1
            continue
        head, i = lines[i], i + 2
        body = []
        while i < len(lines) and not lines[i].strip().startswith("problem main functions"):
            body.append(lines[i])
            i += 1
        vals = []
        for t in re.findall(r"\([^)]*\)|[^\s()]+", "\n".join(body)):
            if t.startswith("("):
                a, b = t[1:-1].replace("\n", "").split(",")
                vals.append(complex(float(a), float(b)))
            else:
                vals.append(complex(float(t), 0.0))
        blocks.append((head, np.array(vals).reshape(-1, 10)))
    return blocks
