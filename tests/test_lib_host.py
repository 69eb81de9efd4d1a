"""The C-ABI library without a GPU: it loads, exports every entry point that
include/movfem_gpu.h declares, rejects bad arguments with error codes (never
aborts), and its host-only zero removal matches find_zeros/rem_zeros."""
import ctypes as C
import pathlib
import re

import numpy as np

from movfem_amd import _lib as L
from movfem_amd import model as M

ROOT = pathlib.Path(__file__).resolve().parent.parent


def header_functions():
    txt = (ROOT / "include" / "movfem_gpu.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mvf_[a-z_0-9]+)\s*\(", txt)))


def test_exports_match_header():
    lib = L.lib()
    names = header_functions()
    assert names, "no declarations found"
    assert set(names) == set(L.EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


def test_bad_arguments_return_codes():
    lib = L.lib()
    h = C.c_void_p()
    assert lib.mvf_init(None, None, None, None, C.byref(h)) == -1
    m = M.random_small(8, 2, 2, 2, seed=1)
    xp, yp, zp = m.xp.copy(), m.yp.copy(), m.zp.copy()
    sg = np.ascontiguousarray(m.sigma).view(np.float64)
    mu = np.ascontiguousarray(m.mu)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    grid = L.Grid(2, 2, 2, 9, dp(xp), dp(yp), dp(zp))  # 9 nodes: invalid
    model = L.Model(dp(sg), dp(mu))
    bc = L.BC(1, 1, 0, 2, 2, 0, 1.0, 1.0, 2.0)
    assert lib.mvf_init(C.byref(grid), C.byref(model), C.byref(bc), None, C.byref(h)) == -1
    grid.nodes_per_elem = 8
    bc.inimod = 4  # Dirichlet boundary models are 1, 2 and 3 (boundary_conds.f90:199-236)
    rc = lib.mvf_init(C.byref(grid), C.byref(model), C.byref(bc), None, C.byref(h))
    assert rc == -1
    assert b"inimod" in lib.mvf_last_error(h)
    lib.mvf_free(h)
    bc.inimod = 3  # layered earth without layers
    bc.nl = 0
    rc = lib.mvf_init(C.byref(grid), C.byref(model), C.byref(bc), None, C.byref(h))
    assert rc == -1
    assert b"inimod 3" in lib.mvf_last_error(h)
    lib.mvf_free(h)
    # no handle: the global_vfem forms refuse before touching a device
    n = C.c_int64(-1)
    i32 = np.zeros(4, np.int32).ctypes.data_as(C.POINTER(C.c_int32))
    f64 = np.zeros(8).ctypes.data_as(C.POINTER(C.c_double))
    assert lib.mvf_global_vfem(None, i32, i32, f64, f64, 4, 2, 2) == -1
    assert lib.mvf_global_vfem_nz(None, i32, i32, f64, f64, 4, 2, 2, C.byref(n)) == -1
    assert n.value == -1


def test_remove_zeros_host():
    lib = L.lib()
    rng = np.random.default_rng(3)
    n = 200
    ia = rng.integers(1, 30, n).astype(np.int32)
    ja = rng.integers(1, 30, n).astype(np.int32)
    a = rng.normal(size=n) + 1j * rng.normal(size=n)
    a[rng.random(n) < 0.3] = 0
    a[5] = 1j * 0.0 + 1e-300  # tiny but non-zero survives
    keep = a != 0
    av = a.copy()
    k = C.c_int64()
    rc = lib.mvf_remove_zeros(n, ia.ctypes.data_as(C.POINTER(C.c_int32)), ja.ctypes.data_as(C.POINTER(C.c_int32)),
                              av.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), C.byref(k))
    assert rc == 0 and k.value == keep.sum()
    assert np.array_equal(av[:k.value], a[keep])
