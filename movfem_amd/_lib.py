"""ctypes bindings of libmovfem_gpu.so (include/movfem_gpu.h)."""
from __future__ import annotations

import ctypes as C
import pathlib

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libmovfem_gpu.so"

MVF_EMU_GPML_PREV_FLAGS = 1
MVF_EMU_UPDATE_SIGMA = 2
MVF_EMU_ALL = 3
MVF_KRYLOV_COCG = 0
MVF_KRYLOV_BICGSTAB = 1
MVF_KRYLOV_GMRES = 2
MVF_PREC_FP64 = 0
MVF_PREC_MIXED = 1
MVF_MATRIX_C64REF = 0
MVF_MATRIX_FP64 = 1
MVF_PC_NONE = 0
MVF_PC_JACOBI = 1
MVF_PC_DIRECT = 2

MVF_OK, MVF_E_ARG, MVF_E_DEVICE, MVF_E_SINGULAR, MVF_E_INDEX = 0, -1, -2, -3, -4
MVF_E_UNSUPPORTED, MVF_E_NOCONV, MVF_E_BREAKDOWN, MVF_E_MEMORY = -5, -6, -7, -8

ERRORS = {0: "MVF_OK", -1: "MVF_E_ARG", -2: "MVF_E_DEVICE", -3: "MVF_E_SINGULAR", -4: "MVF_E_INDEX",
          -5: "MVF_E_UNSUPPORTED", -6: "MVF_E_NOCONV", -7: "MVF_E_BREAKDOWN", -8: "MVF_E_MEMORY"}

# every entry point of include/movfem_gpu.h
EXPORTS = ("mvf_init", "mvf_free", "mvf_last_error", "mvf_sizes", "mvf_get_gne", "mvf_set_frequency",
           "mvf_integrate", "mvf_global_vfem", "mvf_global_vfem_nz", "mvf_remove_zeros", "mvf_default_solver_opts",
           "mvf_solve",
           "mvf_solve_frequencies", "mvf_get_solution", "mvf_get_rhs", "mvf_setup_times", "mvf_assemble",
           "mvf_direct_plan_stats", "mvf_readout", "mvf_write_solution", "mvf_write_receivers", "mvf_write_surface",
           "mvf_read_inputs", "mvf_inputs_view", "mvf_inputs_free", "mvf_grid3d", "mvf_mesh_sizes",
           "mvf_mesh_arrays", "mvf_mesh_free", "mvf_problem_fields", "mvf_write_problem", "mvf_get_operators",
           "mvf_solve_noalloc", "mvf_solve_guess", "mvf_solve_batch", "mvf_get_rows", "mvf_nn_default_config", "mvf_nn_create", "mvf_nn_free",
           "mvf_nn_last_error", "mvf_nn_num_params", "mvf_nn_get_params", "mvf_nn_set_params", "mvf_nn_get_grads",
           "mvf_nn_forward_backward", "mvf_nn_adam", "mvf_nn_predict", "mvf_nn_set_timing", "mvf_nn_stats")


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nodes_per_elem", C.c_int32),
                ("xp", C.POINTER(C.c_double)), ("yp", C.POINTER(C.c_double)), ("zp", C.POINTER(C.c_double))]


class Model(C.Structure):
    _fields_ = [("sigma", C.POINTER(C.c_double)), ("mu", C.POINTER(C.c_double))]


class BC(C.Structure):
    _fields_ = [("dirichlet", C.c_int32), ("inimod", C.c_int32), ("gpml_sch", C.c_int32), ("nextd", C.c_int32),
                ("nzl_top", C.c_int32), ("pad_", C.c_int32), ("a0", C.c_double), ("b0", C.c_double),
                ("nn", C.c_double), ("ztop", C.c_double), ("h_sigma", C.c_double), ("nl", C.c_int32),
                ("pad2_", C.c_int32), ("l_sigma", C.POINTER(C.c_double)), ("l_dz", C.POINTER(C.c_double))]


class Options(C.Structure):
    _fields_ = [("emulate", C.c_int32), ("device", C.c_int32), ("verbose", C.c_int32), ("pad_", C.c_int32)]


class SolverOpts(C.Structure):
    _fields_ = [("method", C.c_int32), ("precision", C.c_int32), ("matrix", C.c_int32), ("precond", C.c_int32),
                ("tol", C.c_double), ("maxit", C.c_int32), ("check_every", C.c_int32), ("inner_tol", C.c_double),
                ("use_graph", C.c_int32), ("gmres_m", C.c_int32), ("freqs_per_factor", C.c_int32),
                ("direct_tile", C.c_int32)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("refinements", C.c_int32), ("converged", C.c_int32),
                ("fallbacks", C.c_int32), ("true_resid", C.c_double * 64), ("seconds", C.c_double),
                ("spmv_launches", C.c_int64 * 2), ("spmv_ms", C.c_double * 2), ("spmv_bytes", C.c_double * 2),
                ("spmv_ncol", C.c_int32), ("spmv_lpr", C.c_int32),
                ("factor_seconds", C.c_double), ("factor_flops", C.c_double), ("factor_bytes", C.c_double),
                ("factorizations", C.c_int32), ("static_pivots", C.c_int32), ("precond_applies", C.c_int32),
                ("operator_fp32", C.c_int32), ("gemm_ms", C.c_double), ("gemm_flops", C.c_double),
                ("gemm_launches", C.c_int64), ("graph_launches", C.c_int64)]


class NnConfig(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("cin", C.c_int32), ("cout", C.c_int32),
                ("width", C.c_int32), ("batch", C.c_int32), ("device", C.c_int32), ("seed", C.c_uint64),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("leaky", C.c_float), ("pad_", C.c_int32)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class MeshInput(C.Structure):
    _fields_ = ([(n, C.c_double) for n in ("dx", "dy", "dz", "high", "depth", "sigma_bg")]
                + [(n, C.c_int32) for n in ("mn", "nsr", "nsf", "mx", "my", "mz", "isigma", "imu", "nf", "pad_")]
                + [("xsr", _dp), ("ysr", _dp), ("zsr", _dp), ("nsp", _ip), ("xto", _dp), ("yto", _dp),
                   ("zto", _dp), ("xm", _dp), ("ym", _dp), ("zm", _dp), ("ijsigma", _ip), ("sigma", _dp),
                   ("ijmu", _ip), ("mu", _dp), ("freqs", _dp)])


class Params(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in ("nf", "boundary", "inimod", "nl", "gpml_sch", "pad_")]
                + [(n, C.c_double) for n in ("h_sigma", "a0", "b0", "nn")]
                + [("freqs", _dp), ("l_sigma", _dp), ("l_dz", _dp)])


_L = None


def lib():
    """Load the HIP library; fails loudly when it has not been built."""
    global _L
    if _L is not None:
        return _L
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m movfem_amd.build`")
    L = C.CDLL(str(LIB_PATH))
    dp, ip, i64p = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    vp = C.c_void_p
    L.mvf_init.restype = C.c_int
    L.mvf_init.argtypes = [C.POINTER(Grid), C.POINTER(Model), C.POINTER(BC), C.POINTER(Options), C.POINTER(vp)]
    L.mvf_free.restype = None
    L.mvf_free.argtypes = [vp]
    L.mvf_last_error.restype = C.c_char_p
    L.mvf_last_error.argtypes = [vp]
    L.mvf_sizes.argtypes = [vp, ip, i64p, i64p]
    L.mvf_get_gne.argtypes = [vp, ip]
    L.mvf_set_frequency.argtypes = [vp, C.c_double]
    L.mvf_integrate.argtypes = [vp, C.c_int64, C.c_int64, dp, dp, dp]
    L.mvf_global_vfem.argtypes = [vp, ip, ip, dp, dp, C.c_int64, C.c_int32, C.c_int32]
    L.mvf_global_vfem_nz.argtypes = [vp, ip, ip, dp, dp, C.c_int64, C.c_int32, C.c_int32, i64p]
    L.mvf_remove_zeros.argtypes = [C.c_int64, ip, ip, dp, i64p]
    L.mvf_default_solver_opts.restype = None
    L.mvf_default_solver_opts.argtypes = [C.POINTER(SolverOpts)]
    L.mvf_solve.argtypes = [vp, dp, C.POINTER(SolverOpts), C.POINTER(SolveInfo)]
    L.mvf_solve_noalloc.argtypes = [vp, dp, C.POINTER(SolverOpts), C.POINTER(SolveInfo)]
    L.mvf_solve_batch.argtypes = [C.POINTER(vp), C.c_int32, dp, C.c_int32, C.POINTER(SolverOpts),
                                  C.POINTER(SolveInfo), ip, C.c_int32]
    L.mvf_solve_guess.argtypes = [vp, dp, dp, C.POINTER(SolverOpts), C.POINTER(SolveInfo), dp]
    fp = C.POINTER(C.c_float)
    L.mvf_nn_default_config.restype = None
    L.mvf_nn_default_config.argtypes = [C.POINTER(NnConfig)]
    L.mvf_nn_create.argtypes = [C.POINTER(NnConfig), C.POINTER(vp)]
    L.mvf_nn_free.restype = None
    L.mvf_nn_free.argtypes = [vp]
    L.mvf_nn_last_error.restype = C.c_char_p
    L.mvf_nn_last_error.argtypes = [vp]
    L.mvf_nn_num_params.restype = C.c_int64
    L.mvf_nn_num_params.argtypes = [vp]
    for fn in (L.mvf_nn_get_params, L.mvf_nn_set_params, L.mvf_nn_get_grads):
        fn.argtypes = [vp, fp]
    L.mvf_nn_forward_backward.argtypes = [vp, C.c_int32, fp, fp, fp, fp]
    L.mvf_nn_adam.argtypes = [vp, C.c_float]
    L.mvf_nn_predict.argtypes = [vp, C.c_int32, fp, fp]
    L.mvf_nn_set_timing.argtypes = [vp, C.c_int32]
    L.mvf_nn_stats.argtypes = [vp, dp]
    L.mvf_solve_frequencies.argtypes = [vp, dp, C.c_int32, C.POINTER(SolverOpts), C.POINTER(SolveInfo)]
    L.mvf_get_solution.argtypes = [vp, dp, C.c_int32]
    L.mvf_get_rhs.argtypes = [vp, dp, C.c_int32]
    L.mvf_setup_times.argtypes = [vp, dp]
    L.mvf_get_operators.argtypes = [vp, C.POINTER(C.c_int64), ip, dp, dp]
    L.mvf_get_rows.argtypes = [vp, C.c_int64, C.c_int64, i64p, ip, dp, dp]
    L.mvf_assemble.argtypes = [vp, C.c_int32, dp]
    L.mvf_direct_plan_stats.argtypes = [C.c_int32] * 4 + [ip, C.c_int64, i64p, ip, dp]
    L.mvf_readout.argtypes = [vp, dp, C.c_int32, C.c_int32, C.c_int32, dp, dp] + [dp] * 11
    cp = C.c_char_p
    L.mvf_write_solution.argtypes = [cp, C.c_int32, C.c_int32, C.c_int64] + [dp] * 5 + [C.c_int64, ip]
    L.mvf_write_receivers.argtypes = [cp] + [C.c_int32] * 4 + [dp] * 8
    L.mvf_write_surface.argtypes = [cp] + [C.c_int32] * 5 + [dp] * 3 + [i64p, C.c_int64] + [dp] * 5
    L.mvf_problem_fields.argtypes = [vp, dp, dp, dp]
    L.mvf_write_problem.argtypes = [cp, C.c_int32, C.c_int32, C.c_int64, dp, dp, dp, C.c_int64, ip]
    L.mvf_read_inputs.argtypes = [cp, cp, C.POINTER(vp)]
    L.mvf_inputs_view.argtypes = [vp, C.POINTER(MeshInput), C.POINTER(Params)]
    L.mvf_inputs_free.restype = None
    L.mvf_inputs_free.argtypes = [vp]
    L.mvf_grid3d.argtypes = [C.POINTER(MeshInput), C.POINTER(vp)]
    L.mvf_mesh_sizes.argtypes = [vp, i64p, dp]
    L.mvf_mesh_arrays.argtypes = [vp] + [dp] * 5 + [ip] + [dp] * 3
    L.mvf_mesh_free.restype = None
    L.mvf_mesh_free.argtypes = [vp]
    for name in EXPORTS:
        if name not in ("mvf_free", "mvf_last_error", "mvf_default_solver_opts", "mvf_inputs_free", "mvf_mesh_free"):
            getattr(L, name).restype = C.c_int
    _L = L
    return L
