"""Synthetic 3DEM_FEM.INP / PARAM.INP files for the reader and grid_3d tests.

The layouts follow the reference's reader (read_input.f90:47-201) and the
PARAM.INP reads of initialize_vfem (MoVFEM_3DMT.f90:302-334).  Every case is
generated from a seed, so the tests and the golden writer see the same text.
Cases vary what the lattice generator branches on: interface relief and point
counts (the Shepard fits), several interfaces, element orders 8/20/27, the
skin-depth branch that sizes the extension zones (depth factor 1.5, 2, 3),
tensor conductivity and permeability components, a receiver whose y lies
outside every other y (the reference takes that receiver's x there), negative
off-diagonal conductivities (which the reference's final pass replaces by the
value of the node below), and the list-directed forms the reader must accept
(tabs, commas, r*c repeats, D exponents, lists continued over records).
"""
from __future__ import annotations

import numpy as np


def _f(v, style):
    s = repr(float(v))
    if style == "d" and "e" in s:
        s = s.replace("e", "D")
    return s


def _join(vals, style):
    if style == "comma":
        return ", ".join(vals)
    if style == "tab":
        return "\t".join(vals)
    return "  ".join(vals)


def render_3dem(c, style="plain"):
    """3DEM_FEM.INP text of case dict c (see make_case)."""
    f = lambda v: _f(v, style)  # noqa: E731
    L = ["-- dx -- dy -- dz -- high -- deph -- nodel-----background_sigma---------",
         _join([f(c["dx"]), f(c["dy"]), f(c["dz"]), f(c["high"]), f(c["depth"]), str(c["mn"]), f(c["sigma_bg"])],
               style),
         "-------------- source & recievers  --------------------------------------",
         str(len(c["sites"]))]
    for i, s in enumerate(c["sites"]):
        L.append(_join([str(i + 1)] + [f(v) for v in s], style))
    L += ["------------------------topography & interfaces -------------------------", f" {len(c['interfaces'])}"]
    counts = [str(len(p)) for p in c["interfaces"]]
    if style == "comma" and len(counts) > 1:  # one list over two records
        L += [", ".join(counts[:1]) + ",", ", ".join(counts[1:])]
    else:
        L.append(_join(counts, style))
    for l, p in enumerate(c["interfaces"]):
        L.append(f"------------------------ interface {l + 1} -----------------------------")
        for j, q in enumerate(p):
            L.append(_join([str(j + 1)] + [f(v) for v in q], style))
    zm = c["zm"]
    mx, my, mz = zm.shape
    ijm = c.get("ijmu") or []
    L += ["----------IEMU---ICDU----MX----NY------mz------------------------",
          _join([str(len(ijm)), str(len(c["ijsigma"])), str(mx), str(my), str(mz)], style)]
    L += [f"MU{a}{b}" for a, b in ijm]
    L += [f"SIGMA{a}{b}" for a, b in c["ijsigma"]]
    L.append("--------------- XM(MX), YM(MY), NZ(MX,MY) -------------------------------")
    for arr in (c["xm"], c["ym"]):
        vals = [f(v) for v in arr]
        if style == "comma":  # one READ over two records
            h = len(vals) // 2
            L += [", ".join(vals[:h]) + ",", ", ".join(vals[h:])]
        else:
            L.append(_join(vals, style))
    L.append("------------- z-coordinates of the 3-D model: zm (mx,my) -------------------------------")
    for i in range(mx):
        for j in range(my):
            L.append(_join([f(v) for v in zm[i, j]], style))

    def block(title, vals):
        L.append(title)
        for i in range(mx):
            for j in range(my):
                row = vals[i, j]
                if style == "comma" and np.all(row == row[0]):
                    L.append(f"{len(row)}*{f(row[0])}")
                else:
                    L.append(_join([f(v) for v in row], style))

    for k, (a, b) in enumerate(c["ijsigma"]):
        block(f"------------------------------ cnd_{a}{b} ----------------------------------", c["sigma"][..., k])
    for k, (a, b) in enumerate(ijm):
        block(f"------------------------------ mu_{a}{b} ----------------------------------", c["mu"][..., k])
    return "\n".join(L) + "\n"


def render_param(freqs, boundary=0, inimod=1, h_sigma=0.01, l_sigma=(0.01,), l_dz=(), gpml_sch=1, a0=1.0, b0=1.0,
                 nn=2.0):
    """PARAM.INP text (the layout of the reference's example file)."""
    L = ["Frequency range [Hz], nf", str(len(freqs)), " ".join(repr(float(v)) for v in freqs),
         "(1) Dirichlet  or (0) GPML  Boundary?", str(boundary),
         "Model for Dirichlet boundary (1) Zero, (2) Homogeneous, (3) Layered", str(inimod),
         "If Homogeneous model, then assign the conductivity of half-space.", repr(float(h_sigma)),
         "If Layered Model, then assign the number of layers (nl)", str(len(l_sigma)),
         "Conductivity of layers (1:nl)", " ".join(repr(float(v)) for v in l_sigma),
         "Thickness of Layers (1:nl-1) (in meters)", " ".join(repr(float(v)) for v in l_dz) or "5000.",
         "PML Scheme: (0) hx=h0x(1+bx/(iwe)) (1) hx=1-ibx/(a0+iw) !If Boundary is PML!", str(gpml_sch),
         "PML Constants: a0,b0,N !If Boundary is PML!", repr(float(a0)), repr(float(b0)), repr(float(nn))]
    return "\n".join(L) + "\n"


def _surface(rng, n, x0, x1, y0, y1, base, amp, kind):
    if kind == "grid":
        k = int(np.ceil(np.sqrt(n)))
        gx, gy = np.meshgrid(np.linspace(x0, x1, k), np.linspace(y0, y1, k), indexing="ij")
        x = gx.ravel()[:n] + rng.uniform(-50, 50, n)
        y = gy.ravel()[:n] + rng.uniform(-50, 50, n)
    else:
        x = rng.uniform(x0, x1, n)
        y = rng.uniform(y0, y1, n)
    z = base + amp * np.sin(x / 3100.0 + 0.3) * np.cos(y / 2600.0) + 0.04 * amp * (x - x0) / 1000.0
    return np.stack([np.round(x, 1), np.round(y, 1), np.round(z, 2)], axis=1)


def make_case(seed, mn=8, ninterf=2, npts=(160, 90, 45), isigma=1, imu=0, sigma_bg=0.1, freqs=(1.0, 10.0),
              dx=1000.0, dy=1000.0, dz=1000.0, negative_offdiag=False, ext=(12000.0, 10000.0)):
    """One synthetic case.  Interfaces deepest first; the last is the surface."""
    rng = np.random.Generator(np.random.PCG64(seed))
    X, Y = ext
    bases = {1: [0.0], 2: [-4000.0, 0.0], 3: [-8000.0, -3000.0, 0.0]}[ninterf]
    amps = {1: [250.0], 2: [500.0, 250.0], 3: [600.0, 400.0, 250.0]}[ninterf]
    interfaces = []
    for l in range(ninterf):
        n = npts[ninterf - 1 - l]
        interfaces.append(_surface(rng, n, -800.0, X + 800.0, -600.0, Y + 600.0, bases[l], amps[l],
                                   "grid" if l % 2 == 0 else "scatter"))
    nsite = 8
    sites = np.stack([np.round(rng.uniform(1500, X - 1500, nsite), 1), np.round(rng.uniform(1500, Y - 1500, nsite), 1),
                      np.zeros(nsite)], axis=1)
    sites[3, 1] = Y + 1900.0  # the largest y of all: the reference takes this site's x for ymax
    mx, my, mz = 7, 6, 5
    xm = np.round(np.linspace(0.0, X, mx), 1)
    ym = np.round(np.linspace(0.0, Y, my), 1)
    zm = np.zeros((mx, my, mz))
    deep = bases[0] - 2500.0
    for i in range(mx):
        for j in range(my):
            zm[i, j] = np.round(np.linspace(deep, -100.0 * (1 + (i + j) % 3), mz), 1)
    comps = {1: [(1, 1)], 3: [(1, 1), (2, 2), (3, 3)],
             6: [(1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]}[isigma]
    sig = np.exp(rng.uniform(np.log(1e-3), np.log(1.0), (mx, my, mz, len(comps))))
    for k, (a, b) in enumerate(comps):
        if a != b:
            sig[..., k] *= 0.1 * (rng.uniform(-1, 1, (mx, my, mz)) if negative_offdiag else 1.0)
    c = dict(dx=dx, dy=dy, dz=dz, high=5000.0, depth=6000.0, mn=mn, sigma_bg=sigma_bg, sites=sites,
             interfaces=interfaces, xm=xm, ym=ym, zm=zm, ijsigma=comps, sigma=np.round(sig, 6), freqs=list(freqs))
    if imu:
        cm = {1: [(1, 1)], 3: [(1, 1), (2, 2), (3, 3)]}[imu]
        c["ijmu"] = cm
        c["mu"] = np.round(rng.uniform(1.0, 1.5, (mx, my, mz, len(cm))), 4)
    return c


# name -> (case arguments, file style); sized so the reference's grid_3d
# finishes in seconds
CASES = {
    "topo8": (dict(seed=101, mn=8, ninterf=2, sigma_bg=0.1, freqs=(1.0, 10.0)), "plain"),
    "three20": (dict(seed=102, mn=20, ninterf=3, isigma=3, imu=1, sigma_bg=0.01, freqs=(0.3, 3.0)), "tab"),
    "tensor27": (dict(seed=103, mn=27, ninterf=2, npts=(140, 60), isigma=6, imu=3, negative_offdiag=True,
                      sigma_bg=0.01, freqs=(0.1,), dx=1200.0, dy=900.0, dz=800.0), "comma"),
    "flat8d": (dict(seed=104, mn=8, ninterf=1, npts=(64,), sigma_bg=0.2, freqs=(2.0, 0.5, 8.0)), "d"),
}


def case_files(name):
    """(3DEM_FEM.INP text, frequencies) of a named case."""
    args, style = CASES[name]
    c = make_case(**args)
    return render_3dem(c, style), c["freqs"]
