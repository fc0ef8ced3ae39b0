"""ctypes binding of liblaamd.so (the C ABI in include/laamd.h).

The shared library is built in-tree by linearalgebra_amd/build.py. Loading
fails loudly when it is missing: there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblaamd.so")

LA_OK, LA_ERR_ARG, LA_ERR_HIP, LA_ERR_ALLOC = 0, 1, 2, 3
LA_ERR_NOT_CONVERGED, LA_ERR_NUMERIC, LA_ERR_STATE, LA_ERR_TIMEOUT = 4, 5, 6, 7
F64, F32, BF16 = 0, 1, 2
SPMV_ROW, SPMV_SEGSUM, SPMV_ALBUS, SPMV_AUTO, SPMV_SELL, SPMV_COLBLOCK = 0, 1, 2, 3, 4, 5
TRSV_LEVEL, TRSV_SYNCFREE, TRSV_JACOBI = 0, 1, 2
LOWER, UPPER = 0, 1
PREC_FP64, PREC_MIXED_F32, PREC_MIXED_BF16, PREC_MIXED_BF16_F32P, PREC_MIXED_F32_BF16P = 0, 1, 2, 3, 4
STOP_TRUE_RESIDUAL, STOP_RECURRENCE = 0, 1


class LaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("laamd error %d: %s" % (code, msg))
        self.code = code


class SolverOpts(C.Structure):
    _fields_ = [("rel_tol", C.c_double), ("abs_tol", C.c_double),
                ("max_iters", C.c_int), ("restart", C.c_int),
                ("stop_rule", C.c_int), ("precision", C.c_int),
                ("trsv_mode", C.c_int), ("spmv_alg", C.c_int),
                ("inner_tol", C.c_double), ("max_outer", C.c_int),
                ("print_level", C.c_int), ("reliable_delta", C.c_double)]


class SolverResult(C.Structure):
    _fields_ = [("converged", C.c_int), ("iters", C.c_int),
                ("outer_iters", C.c_int), ("rel_residual", C.c_double),
                ("seconds", C.c_double)]


_lib = None

# name -> argtypes (restype is int unless listed in _RESTYPES)
_vp, _i, _d, _i64, _u64 = C.c_void_p, C.c_int, C.c_double, C.c_int64, C.c_uint64
_SIGS = {
    "la_last_error": [],
    "la_version": [],
    "la_device_alloc": [C.c_size_t, C.POINTER(_vp)],
    "la_device_free": [_vp],
    "la_memcpy_h2d": [_vp, _vp, C.c_size_t],
    "la_memcpy_d2h": [_vp, _vp, C.c_size_t],
    "la_device_synchronize": [],
    "la_topo_sort2": [_i, _i, _vp, _vp, _i, _vp, _vp, C.POINTER(_i)],
    "la_split_ldu": [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                     C.POINTER(_i), C.POINTER(_i)],
    "la_split_triangle": [_i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp, C.POINTER(_i)],
    "la_diagonal": [_i, _i, _vp, _vp, _vp, _vp, _vp, C.POINTER(_i)],
    "la_transpose": [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "la_permute": [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "la_rcm": [_i, _i, _vp, _vp, _vp, _vp],
    "la_rcm_threads": [_i, _i, _vp, _vp, _i, _vp, _vp],
    "la_iluk_symbolic": [_i, _i, _vp, _vp, _i, _vp, _vp, _vp, C.POINTER(_i64)],
    "la_iluk_numeric": [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "la_icc_symbolic": [_i, _i, _vp, _vp, _vp, _i, _vp, _vp, C.POINTER(_i64)],
    "la_icc_numeric": [_i, _i, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp, C.POINTER(_i)],
    "la_ick_numeric": [_i, _i, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp, C.POINTER(_i)],
    "la_gen_poisson2d": [_i, _i, _vp, _vp, _vp, C.POINTER(_i64)],
    "la_gen_laplace3d": [_i, _i, _i, _vp, _vp, _vp, C.POINTER(_i64)],
    "la_gen_powerlaw": [_i, _d, _d, _i, _u64, _vp, _vp, _vp, C.POINTER(_i64)],
    "la_gen_aniso27": [_i, _i, _i, _d, _u64, _vp, _vp, _vp, C.POINTER(_i64)],
    "la_gen_diffusion2d": [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_i64)],
    "la_gen_uniform": [_i64, _u64, _d, _d, _vp],
    "la_gen_lognormal": [_i64, _u64, _d, _d, _vp],
    "la_mat_create": [_i, _i, _i, _vp, _vp, _vp, C.POINTER(_vp)],
    "la_mat_destroy": [_vp],
    "la_mat_add_dtype": [_vp, _i],
    "la_mat_set_values": [_vp, _vp],
    "la_mat_set_values_device": [_vp, _vp],
    "la_mat_set_stream": [_vp, _vp],
    "la_mat_info": [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i64)],
    "la_spmv_preprocess": [_vp, _i],
    "la_spmv_resolve": [_vp, _i, C.POINTER(_i)],
    "la_mat_sell_info": [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i64)],
    "la_spmv": [_vp, _i, _i, _i, _i, _d, _vp, _d, _vp],
    "la_spmv_axpby": [_vp, _i, _i, _i, _i, _d, _vp, _d, _vp, _vp],
    "la_spmv_transpose": [_vp, _i, _i, _d, _vp, _d, _vp],
    "la_csr_value_grad": [_vp, _vp, _vp, _vp],
    "la_trsv_create": [_i, _i, _vp, _vp, _vp, _vp, _i, C.POINTER(_vp)],
    "la_trsv_destroy": [_vp],
    "la_trsv_add_dtype": [_vp, _i],
    "la_trsv_set_stream": [_vp, _vp],
    "la_trsv_info": [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i64)],
    "la_trsv_set_sweeps": [_vp, _i],
    "la_trsv_solve": [_vp, _i, _i, _i, _vp, _vp],
    "la_trsv_solve_transpose": [_vp, _i, _i, _i, _vp, _vp],
    "la_prec_create_ilu": [_i, _i, _vp, _vp, _vp, C.POINTER(_vp)],
    "la_prec_create_ic": [_i, _i, _vp, _vp, _vp, C.POINTER(_vp)],
    "la_prec_create_ilu_gpu": [_i, _i, _vp, _vp, _i, C.POINTER(_vp)],
    "la_prec_create_ic_gpu": [_i, _i, _vp, _vp, _i, _d, C.POINTER(_vp)],
    "la_prec_create_ick_gpu": [_i, _i, _vp, _vp, _i, _d, C.POINTER(_vp)],
    "la_prec_refactor": [_vp, _vp],
    "la_prec_factor": [_vp, _vp, _vp, _vp, C.POINTER(_i64)],
    "la_prec_destroy": [_vp],
    "la_prec_add_dtype": [_vp, _i],
    "la_prec_set_stream": [_vp, _vp],
    "la_prec_apply": [_vp, _i, _i, _i, _vp, _vp],
    "la_prec_apply_transpose": [_vp, _i, _i, _i, _vp, _vp],
    "la_prec_info": [_vp, C.POINTER(_i), C.POINTER(_i)],
    "la_prec_set_sweeps": [_vp, _i],
    "la_trsv_schedule_info": [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)],
    "la_prec_schedule_info": [_vp, C.POINTER(_i), C.POINTER(_i)],
    "la_dot": [_i64, _i, _vp, _vp, C.POINTER(_d), _vp],
    "la_dot_noalloc": [_i64, _i, _vp, _vp, _vp, _vp],
    "la_axpy_dot": [_i64, _i, _d, _vp, _vp, _vp, C.POINTER(_d), _vp],
    "la_solver_opts_default": [C.POINTER(SolverOpts)],
    "la_pcg_solve": [_vp, _vp, C.POINTER(SolverOpts), _vp, _vp, C.POINTER(SolverResult)],
    "la_fgmres_solve": [_vp, _vp, C.POINTER(SolverOpts), _vp, _vp, C.POINTER(SolverResult)],
    "la_pcg_solve_transpose": [_vp, _vp, C.POINTER(SolverOpts), _vp, _vp,
                               C.POINTER(SolverResult)],
    "la_fgmres_solve_transpose": [_vp, _vp, C.POINTER(SolverOpts), _vp, _vp,
                                  C.POINTER(SolverResult)],
    "la_fgmres_solve_right": [_vp, _vp, _vp, C.POINTER(SolverOpts), _vp, _vp,
                              C.POINTER(SolverResult)],
    "la_rowwise_sqrt_norm": [_i, _i, _vp, _vp, _vp],
    "la_dtad": [_i, _i, _vp, _vp, _vp, _vp],
    "la_mat_rowwise_sqrt_norm": [_vp, _vp],
    "la_mat_dtad": [_vp, _vp],
    "la_read_matrix_market": [C.c_char_p, _i, C.POINTER(_vp)],
    "la_read_binary_coo": [C.c_char_p, _i, C.POINTER(_vp)],
    "la_csr_file_info": [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i64)],
    "la_csr_file_copy": [_vp, _vp, _vp, _vp],
    "la_csr_file_free": [_vp],
    "la_read_binary_vector": [C.c_char_p, _vp, C.POINTER(_i64)],
}
_RESTYPES = {"la_last_error": C.c_char_p, "la_solver_opts_default": None}

EXPORTED = tuple(_SIGS)


def load(path=LIB_PATH):
    """Load liblaamd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("liblaamd.so not built: run `python -m linearalgebra_amd.build` "
                          "(or __graft_entry__.build()) first; no fallback exists")
    # torch carries its own HIP runtime, and the two runtimes in one process
    # share one HSA runtime: torch's must come up first, or this library's
    # later device calls find no device. Bring it up before loading.
    try:
        import torch
        torch.cuda.is_available()
    except ImportError:
        pass
    L = C.CDLL(path)
    for name, args in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _RESTYPES.get(name, C.c_int)
    _lib = L
    return L


def check(status):
    if status != LA_OK:
        msg = load().la_last_error()
        raise LaError(status, msg.decode() if msg else "")
    return status
