"""Run the reference build (_ref/libmovfem_ref.so) on one model, in its own process.

TEST INFRASTRUCTURE ONLY.  The reference's init routines allocate module arrays
without checking, so each model needs a fresh process:

    python -m oracle.refrun in.npz out.npz

in.npz holds the model (see dump_model) and a request list; out.npz receives
gne, nzindx, the element matrices of the requested elements (in request order,
which also advances the GPML flag state) and global_vfem for each frequency.
"""
from __future__ import annotations

import ctypes as C
import subprocess
import sys
import tempfile
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent


def dump_model(m, path, elements=(), vfem_freqs=(), sigma_after=(), readout_x=None, receivers=None,
               write_files=False, write_problem=False):
    rx = np.zeros((0, 2)) if receivers is None else np.asarray(receivers["xy"], np.float64)
    rz = np.zeros(2, np.int32) if receivers is None else np.asarray(receivers["nzl"], np.int32)
    np.savez(path, nx=m.nx, ny=m.ny, nz=m.nz, mn=m.nodes_per_elem, xp=m.xp, yp=m.yp, zp=m.zp,
             sigma=m.sigma, mu=m.mu, freqs=m.freqs, dirichlet=int(m.dirichlet), inimod=m.inimod,
             gpml_sch=m.gpml_sch, a0=m.a0, b0=m.b0, nn=m.nn, nextd=m.nextd, nzl_top=m.nzl_top,
             ztop=m.ztop, h_sigma=m.h_sigma, l_sigma=np.asarray(m.l_sigma, np.float64),
             l_dz=np.asarray(m.l_dz, np.float64), elements=np.asarray(elements, np.int32).reshape(-1, 3),
             vfem_freqs=np.asarray(vfem_freqs, np.int32), sigma_after=np.asarray(sigma_after, np.int32),
             readout_x=np.zeros(0, complex) if readout_x is None else np.asarray(readout_x, np.complex128),
             rx_xy=rx, rx_nzl=rz, write_files=int(write_files), write_problem=int(write_problem))


def parse_receiver_out(path, nsr):
    """Fields, Z, rho_a and phase from the reference's RECEIVER.OUT."""
    lines = pathlib.Path(path).read_text().splitlines()
    num = []
    for ln in lines[2 + nsr:]:
        parts = ln.split()
        try:
            num.append([float(p) for p in parts])
        except ValueError:
            continue  # section header
    fields = np.array(num[:nsr])            # x y, then 12 reals per polarization
    zs = np.array(num[nsr:2 * nsr])         # x y Re/Im of Zxx Zxy Zyx Zyy
    rp = np.array(num[2 * nsr:3 * nsr])     # x y rho(4) phi(4)
    e = np.zeros((nsr, 2, 3), complex)
    hh = np.zeros((nsr, 2, 3), complex)
    for d in range(2):
        v = fields[:, 2 + 12 * d: 2 + 12 * (d + 1)]
        e[:, d, :] = v[:, 0:6:2] + 1j * v[:, 1:6:2]
        hh[:, d, :] = v[:, 6:12:2] + 1j * v[:, 7:12:2]
    z = zs[:, 2::2] + 1j * zs[:, 3::2]
    return {"rx_e": e, "rx_h": hh, "rx_z": z, "rx_rho": rp[:, 2:6], "rx_phi": rp[:, 6:10]}


def run_reference(m, elements=(), vfem_freqs=(), sigma_after=(), timeout=600, readout_x=None, receivers=None,
                  write_files=False, write_problem=False):
    """Returns a dict of reference outputs for model m (fresh subprocess).  write_files:
    also run the reference's SOLUTION.OUT and SURFACE.OUT writers after node_solution;
    their text is returned under "files" with RECEIVER.OUT's.  write_problem: run
    write_problem_ugrid at the first frequency; PROBLEM.OUT's text is returned under
    "problem_out"."""
    import os

    with tempfile.TemporaryDirectory() as td:
        fin, fout = pathlib.Path(td) / "in.npz", pathlib.Path(td) / "out.npz"
        dump_model(m, fin, elements, vfem_freqs, sigma_after, readout_x, receivers, write_files, write_problem)
        env = dict(os.environ)
        env["PYTHONPATH"] = str(HERE.parent) + os.pathsep + env.get("PYTHONPATH", "")
        r = subprocess.run([sys.executable, "-m", "oracle.refrun", str(fin), str(fout)],
                           cwd=td, env=env, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("reference run failed:\n" + r.stdout[-2000:] + r.stderr[-2000:])
        with np.load(fout) as z:
            out = {k: z[k] for k in z.files}
        if receivers is not None:
            out.update(parse_receiver_out(pathlib.Path(td) / "RECEIVER.OUT", len(receivers["xy"])))
        if write_problem:
            out["problem_out"] = (pathlib.Path(td) / "PROBLEM.OUT").read_text()
        if write_files:
            out["files"] = {n: (pathlib.Path(td) / n).read_text()
                            for n in ("SOLUTION.OUT", "RECEIVER.OUT", "SURFACE.OUT") if (pathlib.Path(td) / n).exists()}
        return out


def run_reference_grid(inp_text, freqs, timeout=600):
    """The reference's input() on a 3DEM_FEM.INP holding inp_text, then its
    grid_3d with these frequencies (fresh subprocess).  Returns the reader's
    arrays under in_* and the lattice under g_*."""
    import os

    with tempfile.TemporaryDirectory() as td:
        (pathlib.Path(td) / "3DEM_FEM.INP").write_text(inp_text)
        fin, fout = pathlib.Path(td) / "in.npz", pathlib.Path(td) / "out.npz"
        np.savez(fin, mode="grid", freqs=np.asarray(freqs, np.float64))
        env = dict(os.environ)
        env["PYTHONPATH"] = str(HERE.parent) + os.pathsep + env.get("PYTHONPATH", "")
        r = subprocess.run([sys.executable, "-m", "oracle.refrun", str(fin), str(fout)],
                           cwd=td, env=env, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("reference run failed:\n" + r.stdout[-2000:] + r.stderr[-2000:])
        with np.load(fout) as z:
            return {k: z[k] for k in z.files}


def _grid_main(z, fout, L):
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    L.ref_input_sizes.argtypes = [ip, dp]
    L.ref_input_arrays.argtypes = [dp] * 3 + [ip] + [dp] * 8 + [ip, ip]
    L.ref_grid3d.argtypes = [C.c_int, dp]
    L.ref_grid_sizes.argtypes = [ip, dp]
    L.ref_grid_arrays.argtypes = [dp] * 5 + [ip] + [dp] * 3
    L.ref_read_input()
    isz = np.zeros(9, np.int32)
    rsz = np.zeros(6)
    L.ref_input_sizes(isz.ctypes.data_as(ip), rsz.ctypes.data_as(dp))
    nsr, nsf, nmax, imu, isig, mx, my, mz, mn = (int(v) for v in isz)
    n = mx * my * mz
    out = {"in_isz": isz, "in_rsz": rsz}
    arr = {"in_xsr": np.zeros(nsr), "in_ysr": np.zeros(nsr), "in_zsr": np.zeros(nsr),
           "in_nsp": np.zeros(nsf, np.int32), "in_xto": np.zeros(nsf * nmax), "in_yto": np.zeros(nsf * nmax),
           "in_zto": np.zeros(nsf * nmax), "in_xm": np.zeros(mx), "in_ym": np.zeros(my), "in_zm": np.zeros(n),
           "in_sigma": np.zeros(isig * n), "in_mu": np.zeros(max(imu, 1) * n),
           "in_ijsigma": np.zeros(2 * isig, np.int32), "in_ijmu": np.zeros(max(2 * imu, 1), np.int32)}

    def p(a):
        return a.ctypes.data_as(ip if a.dtype == np.int32 else dp)

    L.ref_input_arrays(*(p(arr[k]) for k in ("in_xsr", "in_ysr", "in_zsr", "in_nsp", "in_xto", "in_yto", "in_zto",
                                              "in_xm", "in_ym", "in_zm", "in_sigma", "in_mu", "in_ijsigma",
                                              "in_ijmu")))
    out.update(arr)
    fr = np.ascontiguousarray(z["freqs"], np.float64)
    L.ref_grid3d(len(fr), fr.ctypes.data_as(dp))
    gsz = np.zeros(11, np.int32)
    gr = np.zeros(5)
    L.ref_grid_sizes(gsz.ctypes.data_as(ip), gr.ctypes.data_as(dp))
    nnx, nny, npt, gnsf, gnsr = int(gsz[3]), int(gsz[4]), int(gsz[6]), int(gsz[8]), int(gsz[10])
    g = {"g_xp": np.zeros(nnx), "g_yp": np.zeros(nny), "g_zp": np.zeros(npt), "g_sigma": np.zeros(12 * npt),
         "g_mu": np.zeros(6 * npt), "g_nzl": np.zeros(gnsf, np.int32), "g_xsr": np.zeros(gnsr),
         "g_ysr": np.zeros(gnsr), "g_zsr": np.zeros(gnsr)}
    L.ref_grid_arrays(*(p(g[k]) for k in ("g_xp", "g_yp", "g_zp", "g_sigma", "g_mu", "g_nzl", "g_xsr", "g_ysr",
                                          "g_zsr")))
    g["g_sigma"] = g["g_sigma"].view(np.complex128).reshape(npt, 6)
    g["g_mu"] = g["g_mu"].reshape(npt, 6)
    out.update(g)
    out["g_isz"], out["g_rsz"] = gsz, gr
    np.savez(fout, **out)


def _main(fin, fout):
    z = np.load(fin)
    L = C.CDLL(str(HERE / "_ref" / "libmovfem_ref.so"))
    if "mode" in z.files and str(z["mode"]) == "grid":
        return _grid_main(z, fout, L)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    L.ref_setup.argtypes = ([C.c_int] * 4 + [dp] * 5 + [C.c_int] * 3 + [C.c_double] * 3 + [C.c_int] * 2
                            + [C.c_double, C.c_int, dp, C.c_double, C.c_int, dp, dp])
    L.ref_set_frequency.argtypes = [C.c_int]
    L.ref_sizes.argtypes = [ip, C.POINTER(C.c_int64)]
    L.ref_gne.argtypes = [ip]
    L.ref_nzindx.argtypes = [ip]
    L.ref_get_sigma.argtypes = [dp]
    L.ref_element.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
    L.ref_global_vfem.argtypes = [ip, ip, dp, dp]
    L.ref_node_solution.argtypes = [dp, dp, dp]
    L.ref_receivers.argtypes = [C.c_int, dp, dp, C.c_int, C.c_int]
    L.ref_write_solution.argtypes = [C.c_int]
    L.ref_write_surface.argtypes = [C.c_int, C.c_int]

    def d(a):
        return a.ctypes.data_as(dp)

    xp, yp, zp = (np.ascontiguousarray(z[k], np.float64) for k in ("xp", "yp", "zp"))
    sg = np.ascontiguousarray(z["sigma"].astype(np.complex128)).view(np.float64)
    mu = np.ascontiguousarray(z["mu"], np.float64)
    fr = np.ascontiguousarray(z["freqs"], np.float64)
    lsg = np.ascontiguousarray(z["l_sigma"], np.float64)
    ldz = np.ascontiguousarray(np.append(z["l_dz"], 0.0), np.float64)
    L.ref_setup(int(z["nx"]), int(z["ny"]), int(z["nz"]), int(z["mn"]), d(xp), d(yp), d(zp), d(sg), d(mu),
                int(z["dirichlet"]), int(z["inimod"]), int(z["gpml_sch"]), float(z["a0"]), float(z["b0"]),
                float(z["nn"]), int(z["nextd"]), int(z["nzl_top"]), float(z["ztop"]), len(fr), d(fr),
                float(z["h_sigma"]), len(lsg), d(lsg), d(ldz))
    nne = C.c_int32()
    nnze = C.c_int64()
    L.ref_sizes(C.byref(nne), C.byref(nnze))
    mn = int(z["mn"])
    me = {8: 12, 20: 36, 27: 54}[mn]
    ne = int(z["nx"]) * int(z["ny"]) * int(z["nz"])
    out = {"nne": np.int64(nne.value), "nnze": np.int64(nnze.value)}
    g = np.empty(ne * me, np.int32)
    L.ref_gne(g.ctypes.data_as(ip))
    out["gne"] = g.reshape(ne, me)
    nz = np.empty(ne * me * me, np.int32)
    L.ref_nzindx(nz.ctypes.data_as(ip))
    out["nzindx"] = nz.reshape(ne, me, me)
    npt = len(zp)
    L.ref_set_frequency(1)
    if "write_problem" in z.files and int(z["write_problem"]):
        L.ref_write_problem()
    els = z["elements"]
    if len(els):
        A = np.empty((len(els), me * me * 2))
        B = np.empty((len(els), me * 2 * 2))
        for t, (ie, je, ke) in enumerate(els):
            a = np.empty(me * me * 2)
            b = np.empty(me * 4)
            L.ref_element(int(ie), int(je), int(ke), d(a), d(b))
            A[t], B[t] = a, b
        out["aloc"] = A.view(np.complex128).reshape(len(els), me, me)
        out["bloc"] = B.view(np.complex128).reshape(len(els), me, 2)
    for fi in z["vfem_freqs"]:
        fi = int(fi)
        if fi != 1:
            L.ref_set_frequency(fi)
        L.ref_sizes(C.byref(nne), C.byref(nnze))
        n = nnze.value
        ia = np.empty(n, np.int32)
        ja = np.empty(n, np.int32)
        a = np.empty(2 * n)
        b = np.empty(4 * nne.value)
        L.ref_global_vfem(ia.ctypes.data_as(ip), ja.ctypes.data_as(ip), d(a), d(b))
        out[f"ia_{fi}"], out[f"ja_{fi}"] = ia, ja
        out[f"a_{fi}"], out[f"b_{fi}"] = a.view(np.complex128), b.view(np.complex128)
    for fi in z["sigma_after"]:
        L.ref_set_frequency(int(fi))
        s = np.empty(npt * 12)
        L.ref_get_sigma(d(s))
        out[f"sigma_{int(fi)}"] = s.view(np.complex128).reshape(npt, 6)
    x = z["readout_x"]
    if len(x):
        xv = np.ascontiguousarray(x, np.complex128).view(np.float64)
        eo = np.empty(2 * npt * 3 * 2)
        ho = np.empty(2 * npt * 3 * 2)
        L.ref_node_solution(d(xv), d(eo), d(ho))
        # Fortran (2*npt, 3) column-major -> (2, npt, 3)
        out["esol"] = eo.view(np.complex128).reshape(3, 2, npt).transpose(1, 2, 0)
        out["hsol"] = ho.view(np.complex128).reshape(3, 2, npt).transpose(1, 2, 0)
        rxy = z["rx_xy"]
        if int(z["write_files"]):
            L.ref_write_solution(0)
        if len(rxy):
            xs = np.ascontiguousarray(rxy[:, 0])
            ys = np.ascontiguousarray(rxy[:, 1])
            L.ref_receivers(len(xs), d(xs), d(ys), int(z["rx_nzl"][0]), int(z["rx_nzl"][1]))
        if int(z["write_files"]):
            L.ref_write_surface(int(z["rx_nzl"][0]), int(z["rx_nzl"][1]))
    np.savez(fout, **out)


if __name__ == "__main__":
    _main(sys.argv[1], sys.argv[2])
