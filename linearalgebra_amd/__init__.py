"""MI355X-native sparse hot path: CSR SpMV, ILU(k)/IC(k) apply over
level-scheduled / sync-free triangular solves, PCG and FGMRES, FP64 and
mixed FP32/BF16 precision.

All work goes through liblaamd.so (HIP kernels for gfx950 + host C++ behind
the C ABI of include/laamd.h). This module is a thin Python view of that ABI
used by the tests, the benchmark and the autograd test binding; device
vectors are passed as torch tensors (torch only provides device memory and
streams). Host-side functions take numpy arrays.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import (BF16, F32, F64, LOWER, PREC_FP64, PREC_MIXED_BF16,
                   PREC_MIXED_BF16_F32P, PREC_MIXED_F32, PREC_MIXED_F32_BF16P, SPMV_ALBUS, SPMV_AUTO, SPMV_COLBLOCK, SPMV_ROW, SPMV_SELL,
                   SPMV_SEGSUM, STOP_RECURRENCE, STOP_TRUE_RESIDUAL,
                   TRSV_JACOBI, TRSV_LEVEL, TRSV_SYNCFREE, UPPER, LaError, SolverOpts,
                   SolverResult, check)

__all__ = [
    "CSR", "DeviceMatrix", "TriangularSolve", "Preconditioner", "pcg",
    "fgmres", "topo_sort2", "split_ldu", "split_triangle", "diagonal",
    "transpose", "permute", "rcm", "iluk_symbolic", "iluk_numeric",
    "icc_symbolic", "icc_numeric", "ick_numeric", "ick_shifted", "gen", "F64", "F32", "BF16",
    "LaError",
    "read_matrix_market", "read_binary_coo", "read_binary_vector", "rowwise_sqrt_norm", "dtad",
]


def lib():
    return _lib.load()


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _dptr(t, dtype_code=None):
    """Device pointer of a contiguous CUDA/HIP torch tensor."""
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    if dtype_code is not None:
        import torch
        want = {F64: torch.float64, F32: torch.float32}[dtype_code]
        if t.dtype != want:
            raise TypeError("expected %s, got %s" % (want, t.dtype))
    return C.c_void_p(t.data_ptr())


def _vec_code(t):
    import torch
    if t.dtype == torch.float64:
        return F64
    if t.dtype == torch.float32:
        return F32
    raise TypeError("vectors must be float64 or float32")


class CSR:
    """Host CSR (int32 indices, float64 values, base = ai[0]); the
    CSRMatrix of sparse_mat_op/matrix_utils.hpp:58-103."""

    def __init__(self, n, ai, aj, av, ncols=None):
        self.n = int(n)
        self.ncols = int(n if ncols is None else ncols)
        self.ai, self.aj, self.av = _i32(ai), _i32(aj), _f64(av)

    @property
    def base(self):
        return int(self.ai[0])

    @property
    def nnz(self):
        return int(self.ai[self.n] - self.ai[0])


# ---------------------------------------------------------------- device


class DeviceMatrix:
    """la_mat_t: an uploaded CSR matrix for SpMV (SPMV<CSR,Impl>,
    sparse_mat_op/spmv.hpp:310-334)."""

    def __init__(self, A, stream=None):
        h = C.c_void_p()
        check(lib().la_mat_create(A.n, A.ncols, A.base, _p(A.ai), _p(A.aj),
                                  _p(A.av), C.byref(h)))
        self.h, self.n, self.ncols, self.nnz = h, A.n, A.ncols, A.nnz
        if stream is not None:
            self.set_stream(stream)

    def __del__(self):
        try:
            if self.h:
                lib().la_mat_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_stream(self, stream):
        ptr = stream if isinstance(stream, int) else stream.cuda_stream
        check(lib().la_mat_set_stream(self.h, C.c_void_p(ptr)))

    def add_dtype(self, dt):
        check(lib().la_mat_add_dtype(self.h, dt))

    def set_values(self, av):
        if hasattr(av, "is_cuda"):
            check(lib().la_mat_set_values_device(self.h, _dptr(av, F64)))
        else:
            check(lib().la_mat_set_values(self.h, _p(_f64(av))))

    def preprocess(self, alg=SPMV_ALBUS):
        check(lib().la_spmv_preprocess(self.h, alg))

    def sell_info(self):
        """{"built", "col16", "slots"} of the sliced-ELL copy (la_mat_sell_info)."""
        b, c16, sl = C.c_int(), C.c_int(), C.c_int64()
        check(lib().la_mat_sell_info(self.h, C.byref(b), C.byref(c16), C.byref(sl)))
        return {"built": bool(b.value), "col16": bool(c16.value), "slots": sl.value}

    def resolve(self, alg=SPMV_AUTO):
        """The kernel family `alg` runs on for this matrix (SPMV_* code)."""
        out = C.c_int(0)
        check(lib().la_spmv_resolve(self.h, alg, C.byref(out)))
        return out.value

    def spmv(self, x, y, alpha=1.0, beta=0.0, alg=SPMV_AUTO, val=None, acc=None, z=None):
        """y = alpha*A*x + beta*(z if given else y)."""
        vec = _vec_code(x)
        if _vec_code(y) != vec:
            raise TypeError("x and y must share a dtype")
        val = vec if val is None else val
        if acc is None:
            acc = F64 if (val == F64 and vec == F64) else F32
        if z is None:
            check(lib().la_spmv(self.h, alg, val, vec, acc, alpha, _dptr(x), beta, _dptr(y)))
        else:
            check(lib().la_spmv_axpby(self.h, alg, val, vec, acc, alpha, _dptr(x), beta,
                                      _dptr(z), _dptr(y)))
        return y

    def spmv_transpose(self, x, y, alpha=1.0, beta=0.0, val=None):
        vec = _vec_code(x)
        val = vec if val is None else val
        check(lib().la_spmv_transpose(self.h, val, vec, alpha, _dptr(x), beta, _dptr(y)))
        return y

    def value_grad(self, lam, x, g):
        check(lib().la_csr_value_grad(self.h, _dptr(lam, F64), _dptr(x, F64), _dptr(g, F64)))
        return g

    def rowwise_sqrt_norm(self, d):
        """rowwiseSqrtNorm of the device values into the device tensor d."""
        check(lib().la_mat_rowwise_sqrt_norm(self.h, _dptr(d, F64)))
        return d

    def dtad(self, d):
        """DtAD in place on the device values (d: device tensor)."""
        check(lib().la_mat_dtad(self.h, _dptr(d, F64)))
        return self


class TriangularSolve:
    """la_trsv_t: OptimizedTriangularSolve (triangle_solve.hpp:147-208) for a
    strict triangle T plus an optional diagonal (None = unit)."""

    def __init__(self, T, diag=None, upper=False):
        h = C.c_void_p()
        d = None if diag is None else _f64(diag)
        check(lib().la_trsv_create(T.n, T.base, _p(T.ai), _p(T.aj), _p(T.av), _p(d),
                                   UPPER if upper else LOWER, C.byref(h)))
        self.h, self.n = h, T.n

    def __del__(self):
        try:
            if self.h:
                lib().la_trsv_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def info(self):
        n, lv, nnz = C.c_int(), C.c_int(), C.c_int64()
        check(lib().la_trsv_info(self.h, C.byref(n), C.byref(lv), C.byref(nnz)))
        t, parts, edge = C.c_int(), C.c_int(), C.c_int()
        check(lib().la_trsv_schedule_info(self.h, C.byref(t), C.byref(parts), C.byref(edge)))
        return {"n": n.value, "levels": lv.value, "nnz": nnz.value, "tiled": bool(t.value),
                "tiles": parts.value, "tile_edge": edge.value}

    def add_dtype(self, dt):
        check(lib().la_trsv_add_dtype(self.h, dt))

    def set_sweeps(self, k):
        """Sweeps of TRSV_JACOBI solves (la_trsv_set_sweeps)."""
        check(lib().la_trsv_set_sweeps(self.h, int(k)))
        return self

    def solve(self, b, x, mode=TRSV_SYNCFREE, val=None):
        vec = _vec_code(b)
        val = vec if val is None else val
        check(lib().la_trsv_solve(self.h, mode, val, vec, _dptr(b), _dptr(x)))
        return x

    def solve_transpose(self, b, x, mode=TRSV_SYNCFREE, val=None):
        vec = _vec_code(b)
        val = vec if val is None else val
        check(lib().la_trsv_solve_transpose(self.h, mode, val, vec, _dptr(b), _dptr(x)))
        return x


class Preconditioner:
    """la_prec_t: ILU(k) (combined LU factor) or IC(k) (upper factor R,
    M = R^T R) apply, incomplete_lu.cpp:36-48 / incomplete_cholesky.cpp:21-29."""

    def __init__(self, F, kind, _handle=None):
        if _handle is not None:
            self.h, self.n, self.kind = _handle, F, kind
            return
        h = C.c_void_p()
        fn = {"ilu": lib().la_prec_create_ilu, "ic": lib().la_prec_create_ic}[kind]
        check(fn(F.n, F.base, _p(F.ai), _p(F.aj), _p(F.av), C.byref(h)))
        self.h, self.n, self.kind = h, F.n, kind

    @classmethod
    def ilu_gpu(cls, A, level):
        """ILU(level) of A with the numeric phase on the device: call
        refactor(values) before use (la_prec_create_ilu_gpu)."""
        h = C.c_void_p()
        check(lib().la_prec_create_ilu_gpu(A.n, A.base, _p(A.ai), _p(A.aj), level, C.byref(h)))
        return cls(A.n, "ilu", _handle=h)

    @classmethod
    def ic_gpu(cls, A, level, omega=0.0):
        """IC(level) of the symmetric A (its upper triangle), numeric phase on
        the device (la_prec_create_ic_gpu)."""
        h = C.c_void_p()
        check(lib().la_prec_create_ic_gpu(A.n, A.base, _p(A.ai), _p(A.aj), level, float(omega),
                                          C.byref(h)))
        return cls(A.n, "ic", _handle=h)

    @classmethod
    def ick_gpu(cls, A, level, shift=0.0):
        """IC(level) of the symmetric A with incomplete_cholesky_k's numeric
        arithmetic (additive shift, first-in first-out rows, a pivot <= 0
        fails; incomplete_cholesky.cpp:195-274), on the device
        (la_prec_create_ick_gpu)."""
        h = C.c_void_p()
        check(lib().la_prec_create_ick_gpu(A.n, A.base, _p(A.ai), _p(A.aj), level, float(shift),
                                           C.byref(h)))
        return cls(A.n, "ic", _handle=h)

    def refactor(self, a_vals):
        """Numeric factorisation on the device from A's FP64 values (a
        device tensor in the CSR order given at creation)."""
        check(lib().la_prec_refactor(self.h, _dptr(a_vals, F64)))
        return self

    def factor(self):
        """The factor (ILU: combined LU; IC: upper R), 0-based host CSR."""
        nnz = C.c_int64()
        check(lib().la_prec_factor(self.h, None, None, None, C.byref(nnz)))
        fai = np.zeros(self.n + 1, np.int32)
        faj = np.zeros(nnz.value, np.int32)
        fav = np.zeros(nnz.value)
        check(lib().la_prec_factor(self.h, _p(fai), _p(faj), _p(fav), C.byref(nnz)))
        return CSR(self.n, fai, faj, fav)

    def __del__(self):
        try:
            if self.h:
                lib().la_prec_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def levels(self):
        a, b = C.c_int(), C.c_int()
        check(lib().la_prec_info(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def tiles(self):
        """Tile counts of the lower and upper solve's tiled sync-free
        schedule (0: the level-chunk schedule)."""
        a, b = C.c_int(), C.c_int()
        check(lib().la_prec_schedule_info(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def add_dtype(self, dt):
        check(lib().la_prec_add_dtype(self.h, dt))

    def set_sweeps(self, k):
        """Sweeps per triangle of TRSV_JACOBI applies (la_prec_set_sweeps)."""
        check(lib().la_prec_set_sweeps(self.h, int(k)))
        return self

    def set_stream(self, stream):
        ptr = stream if isinstance(stream, int) else stream.cuda_stream
        check(lib().la_prec_set_stream(self.h, C.c_void_p(ptr)))

    def apply(self, r, z, mode=TRSV_SYNCFREE, val=None):
        vec = _vec_code(r)
        val = vec if val is None else val
        check(lib().la_prec_apply(self.h, mode, val, vec, _dptr(r), _dptr(z)))
        return z

    def apply_transpose(self, r, z, mode=TRSV_SYNCFREE, val=None):
        vec = _vec_code(r)
        val = vec if val is None else val
        check(lib().la_prec_apply_transpose(self.h, mode, val, vec, _dptr(r), _dptr(z)))
        return z


def solver_opts(**kw):
    o = SolverOpts()
    lib().la_solver_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def _solve(fn, A, P, b, x, raise_on_fail, kw):
    o = solver_opts(**kw)
    r = SolverResult()
    st = fn(A.h, None if P is None else P.h, C.byref(o), _dptr(b, F64), _dptr(x, F64),
            C.byref(r))
    if st != 0 and (raise_on_fail or st != _lib.LA_ERR_NOT_CONVERGED):
        check(st)
    return {"converged": bool(r.converged), "iters": r.iters, "outer_iters": r.outer_iters,
            "rel_residual": r.rel_residual, "seconds": r.seconds, "status": st}


def pcg(A, P, b, x, raise_on_fail=False, **kw):
    """mkl_pcg_solver::solve (mkl_wrapper/mkl_solver.cpp:150-251)."""
    return _solve(lib().la_pcg_solve, A, P, b, x, raise_on_fail, kw)


def fgmres(A, P, b, x, raise_on_fail=False, **kw):
    """mkl_fgmres_solver::solve (mkl_wrapper/mkl_solver.cpp:253-386)."""
    return _solve(lib().la_fgmres_solve, A, P, b, x, raise_on_fail, kw)


def pcg_transpose(A, P, b, x, raise_on_fail=False, **kw):
    """A^T x = b (transposed product and preconditioner apply)."""
    return _solve(lib().la_pcg_solve_transpose, A, P, b, x, raise_on_fail, kw)


def fgmres_right(A, P, R, b, x, raise_on_fail=False, **kw):
    """mkl_fgmres_solver(A, P, R)::solve (mkl_solver.cpp:253-386 with the
    right operator of :319-324): x <- u with A R^-1 u = b."""
    o = solver_opts(**kw)
    r = SolverResult()
    st = lib().la_fgmres_solve_right(A.h, None if P is None else P.h, R.h, C.byref(o),
                                     _dptr(b, F64), _dptr(x, F64), C.byref(r))
    if st != 0 and (raise_on_fail or st != _lib.LA_ERR_NOT_CONVERGED):
        check(st)
    return {"converged": bool(r.converged), "iters": r.iters, "outer_iters": r.outer_iters,
            "rel_residual": r.rel_residual, "seconds": r.seconds, "status": st}


def fgmres_transpose(A, P, b, x, raise_on_fail=False, **kw):
    return _solve(lib().la_fgmres_solve_transpose, A, P, b, x, raise_on_fail, kw)


# ---------------------------------------------------------------- scaling


def rowwise_sqrt_norm(A):
    """mkl_sparse_mat::rowwiseSqrtNorm (mkl_sparse_mat.cpp:517-533)."""
    d = np.zeros(A.n)
    check(lib().la_rowwise_sqrt_norm(A.n, A.base, _p(A.ai), _p(A.av), _p(d)))
    return d


def dtad(A, d):
    """mkl_sparse_mat::DtAD (mkl_sparse_mat.cpp:535-544) -> a new CSR."""
    av = np.array(A.av, dtype=np.float64)
    check(lib().la_dtad(A.n, A.base, _p(A.ai), _p(A.aj), _p(av), _p(_f64(d))))
    return CSR(A.n, A.ai, A.aj, av, ncols=A.ncols)


# ---------------------------------------------------------------- readers


def _read_csr(fn, path, base):
    h = C.c_void_p()
    check(fn(os.fsencode(path), base, C.byref(h)))
    try:
        r, c, nnz = C.c_int(), C.c_int(), C.c_int64()
        check(lib().la_csr_file_info(h, C.byref(r), C.byref(c), C.byref(nnz)))
        ai = np.zeros(r.value + 1, np.int32)
        aj = np.zeros(nnz.value, np.int32)
        av = np.zeros(nnz.value)
        check(lib().la_csr_file_copy(h, _p(ai), _p(aj), _p(av)))
    finally:
        lib().la_csr_file_free(h)
    return CSR(r.value, ai, aj, av, ncols=c.value)


def read_matrix_market(path, base=0):
    """Matrix Market coordinate file -> CSR sorted by (row, col), symmetric
    files expanded (read_matrix_market_csr, utils/utils.h:93-132)."""
    return _read_csr(lib().la_read_matrix_market, path, base)


def read_binary_coo(path, base=1):
    """The reference's binary COO file -> CSR in `base`
    (ReadFromBinaryCSR, utils/utils.h:24-91)."""
    return _read_csr(lib().la_read_binary_coo, path, base)


def read_binary_vector(path):
    """ReadFromBinaryVec (utils/utils.cpp:36-46)."""
    n = C.c_int64()
    check(lib().la_read_binary_vector(os.fsencode(path), None, C.byref(n)))
    out = np.zeros(n.value)
    check(lib().la_read_binary_vector(os.fsencode(path), _p(out), C.byref(n)))
    return out


# ---------------------------------------------------------------- host


def topo_sort2(T, upper):
    iperm = np.zeros(max(T.n, 1), np.int32)
    prefix = np.zeros(T.n + 2, np.int32)
    lv = C.c_int()
    check(lib().la_topo_sort2(T.n, T.base, _p(T.ai), _p(T.aj), 1 if upper else 0,
                              _p(iperm), _p(prefix), C.byref(lv)))
    return iperm[:T.n], prefix[:lv.value + 1], lv.value


def split_ldu(A):
    n = A.n
    Lai, Uai = np.zeros(n + 1, np.int32), np.zeros(n + 1, np.int32)
    ln, un = C.c_int(), C.c_int()
    check(lib().la_split_ldu(n, A.base, _p(A.ai), _p(A.aj), _p(A.av), _p(Lai), None, None,
                             None, _p(Uai), None, None, C.byref(ln), C.byref(un)))
    Laj, Lav = np.zeros(ln.value, np.int32), np.zeros(ln.value)
    Uaj, Uav = np.zeros(un.value, np.int32), np.zeros(un.value)
    D = np.zeros(n)
    check(lib().la_split_ldu(n, A.base, _p(A.ai), _p(A.aj), _p(A.av), _p(Lai), _p(Laj),
                             _p(Lav), _p(D), _p(Uai), _p(Uaj), _p(Uav), C.byref(ln),
                             C.byref(un)))
    return CSR(n, Lai, Laj, Lav), D, CSR(n, Uai, Uaj, Uav)


def split_triangle(A, upper=True):
    n = A.n
    Tai = np.zeros(n + 1, np.int32)
    t = C.c_int()
    check(lib().la_split_triangle(n, A.base, _p(A.ai), _p(A.aj), _p(A.av), 1 if upper else 0,
                                  _p(Tai), None, None, C.byref(t)))
    Taj, Tav = np.zeros(t.value, np.int32), np.zeros(t.value)
    check(lib().la_split_triangle(n, A.base, _p(A.ai), _p(A.aj), _p(A.av), 1 if upper else 0,
                                  _p(Tai), _p(Taj), _p(Tav), C.byref(t)))
    return CSR(n, Tai, Taj, Tav)


def diagonal(A):
    pos, d = np.zeros(A.n, np.int32), np.zeros(A.n)
    ok = C.c_int()
    check(lib().la_diagonal(A.n, A.base, _p(A.ai), _p(A.aj), _p(A.av), _p(pos), _p(d),
                            C.byref(ok)))
    return bool(ok.value), pos, d


def transpose(A):
    ti = np.zeros(A.ncols + 1, np.int32)
    tj, tv = np.zeros(A.nnz, np.int32), np.zeros(A.nnz)
    check(lib().la_transpose(A.n, A.ncols, A.base, _p(A.ai), _p(A.aj), _p(A.av), _p(ti),
                             _p(tj), _p(tv)))
    return CSR(A.ncols, ti, tj, tv, ncols=A.n)


def permute(A, iperm, perm):
    pai = np.zeros(A.n + 1, np.int32)
    paj, pav = np.zeros(A.nnz, np.int32), np.zeros(A.nnz)
    ip = None if iperm is None else _i32(iperm)
    pp = None if perm is None else _i32(perm)
    check(lib().la_permute(A.n, A.base, _p(A.ai), _p(A.aj), _p(A.av), _p(ip), _p(pp),
                           _p(pai), _p(paj), _p(pav)))
    return CSR(A.n, pai, paj, pav)


def rcm(A, nthreads=0):
    """Reverse Cuthill-McKee (reordering/Reordering.cpp:37-136), bit-exact
    with the serial definition for any thread count (0: the default)."""
    iperm, perm = np.zeros(A.n, np.int32), np.zeros(A.n, np.int32)
    check(lib().la_rcm_threads(A.n, A.base, _p(A.ai), _p(A.aj), int(nthreads), _p(iperm),
                               _p(perm)))
    return iperm, perm


def iluk_symbolic(A, level):
    n = A.n
    fai, fu = np.zeros(n + 1, np.int32), np.zeros(max(n, 1), np.int32)
    nnz = C.c_int64()
    check(lib().la_iluk_symbolic(n, A.base, _p(A.ai), _p(A.aj), level, _p(fai), None,
                                 _p(fu), C.byref(nnz)))
    faj = np.zeros(nnz.value, np.int32)
    check(lib().la_iluk_symbolic(n, A.base, _p(A.ai), _p(A.aj), level, _p(fai), _p(faj),
                                 _p(fu), C.byref(nnz)))
    return fai, faj, fu[:n]


def iluk_numeric(A, fai, faj, fu):
    fav = np.zeros(len(faj))
    check(lib().la_iluk_numeric(A.n, A.base, _p(A.ai), _p(A.aj), _p(A.av), _p(_i32(fai)),
                                _p(_i32(faj)), _p(_i32(fu)), _p(fav)))
    return CSR(A.n, fai, faj, fav)


def iluk(A, level):
    fai, faj, fu = iluk_symbolic(A, level)
    return iluk_numeric(A, fai, faj, fu)


def icc_symbolic(U, diag_pos, lvl):
    n = U.n
    ia = np.zeros(n + 1, np.int32)
    nnz = C.c_int64()
    dp = _i32(diag_pos)
    check(lib().la_icc_symbolic(n, U.base, _p(U.ai), _p(U.aj), _p(dp), lvl, _p(ia), None,
                                C.byref(nnz)))
    ja = np.zeros(nnz.value, np.int32)
    check(lib().la_icc_symbolic(n, U.base, _p(U.ai), _p(U.aj), _p(dp), lvl, _p(ia), _p(ja),
                                C.byref(nnz)))
    return ia, ja


def icc_numeric(U, diag_pos, ia, ja, omega=0.0):
    va = np.zeros(len(ja))
    ok = C.c_int()
    check(lib().la_icc_numeric(U.n, U.base, _p(U.ai), _p(U.aj), _p(U.av), _p(_i32(diag_pos)),
                               omega, _p(_i32(ia)), _p(_i32(ja)), _p(va), C.byref(ok)))
    return CSR(U.n, ia, ja, va), bool(ok.value)


def icc(U, lvl, omega=0.0, diag_pos=None):
    """IC(k) of the upper triangle U incl. diagonal -> (R, ok), A ~ R^T R."""
    dp = U.ai[:-1] if diag_pos is None else diag_pos
    ia, ja = icc_symbolic(U, dp, lvl)
    return icc_numeric(U, dp, ia, ja, omega)


def icc_shifted(U, lvl, initial_shift=1e-3, nrestart=20):
    """incomplete_cholesky_k_2::numeric_factorize retry loop
    (mkl_wrapper/incomplete_cholesky.cpp:315-342): multiplicative diagonal
    shift doubled from initial_shift until no negative pivot."""
    dp = U.ai[:-1]
    ia, ja = icc_symbolic(U, dp, lvl)
    shift, it = 0.0, 0
    while True:
        R, ok = icc_numeric(U, dp, ia, ja, shift)
        if ok:
            return R, shift
        it += 1
        if it > nrestart:
            raise LaError(_lib.LA_ERR_NUMERIC, "IC(k) failed after %d shifts" % nrestart)
        shift = max(2 * shift, initial_shift)


def ick_numeric(U, diag_pos, ia, ja, shift=0.0):
    """One pass of incomplete_cholesky_k::numeric_factorize
    (mkl_wrapper/incomplete_cholesky.cpp:195-274): additive diagonal shift,
    first-in first-out row order, fails on a pivot <= 0 (la_ick_numeric)."""
    va = np.zeros(len(ja))
    ok = C.c_int()
    check(lib().la_ick_numeric(U.n, U.base, _p(U.ai), _p(U.aj), _p(U.av), _p(_i32(diag_pos)),
                               float(shift), _p(_i32(ia)), _p(_i32(ja)), _p(va), C.byref(ok)))
    return CSR(U.n, ia, ja, va), bool(ok.value)


def ick_shifted(U, lvl, shift=True, initial_shift=1e-3, nrestart=20):
    """incomplete_cholesky_k symbolic + numeric with its retry loop
    (mkl_wrapper/incomplete_cholesky.cpp:195-274) on the upper triangle U
    (diagonal first in each row) -> (R, shift used). With shift=True the
    first shift is initial_shift - min a_ii when min a_ii <= 0 (else 0), and
    each failure sets shift = max(initial_shift, 2 shift), at most nrestart
    times; with shift=False a pivot <= 0 raises at once."""
    dp = U.ai[:-1]
    ia, ja = icc_symbolic(U, dp, lvl)
    s = 0.0
    if shift and U.n:
        min_diag = float(np.min(np.asarray(U.av)[np.asarray(dp) - U.base]))
        if min_diag <= 0:
            s = initial_shift - min_diag
    it = 0
    while True:
        R, ok = ick_numeric(U, dp, ia, ja, s)
        if ok:
            return R, s
        it += 1
        if not shift or it > nrestart:
            raise LaError(_lib.LA_ERR_NUMERIC, "incomplete_cholesky_k: non-positive pivot")
        s = max(initial_shift, 2 * s)


# ---------------------------------------------------------------- generators


class gen:
    """Seeded synthetic inputs (host, deterministic for a seed)."""

    @staticmethod
    def poisson2d(nx, ny=None):
        ny = nx if ny is None else ny
        n = nx * ny
        ai = np.zeros(n + 1, np.int32)
        nnz = C.c_int64()
        check(lib().la_gen_poisson2d(nx, ny, _p(ai), None, None, C.byref(nnz)))
        aj, av = np.zeros(nnz.value, np.int32), np.zeros(nnz.value)
        check(lib().la_gen_poisson2d(nx, ny, _p(ai), _p(aj), _p(av), C.byref(nnz)))
        return CSR(n, ai, aj, av)

    @staticmethod
    def laplace3d(nx, ny=None, nz=None):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        n = nx * ny * nz
        ai = np.zeros(n + 1, np.int32)
        nnz = C.c_int64()
        check(lib().la_gen_laplace3d(nx, ny, nz, _p(ai), None, None, C.byref(nnz)))
        aj, av = np.zeros(nnz.value, np.int32), np.zeros(nnz.value)
        check(lib().la_gen_laplace3d(nx, ny, nz, _p(ai), _p(aj), _p(av), C.byref(nnz)))
        return CSR(n, ai, aj, av)

    @staticmethod
    def powerlaw(n, mean_len=16.0, sigma=1.0, max_len=20000, seed=3):
        ai = np.zeros(n + 1, np.int32)
        nnz = C.c_int64()
        check(lib().la_gen_powerlaw(n, mean_len, sigma, max_len, seed, _p(ai), None, None,
                                    C.byref(nnz)))
        aj, av = np.zeros(nnz.value, np.int32), np.zeros(nnz.value)
        check(lib().la_gen_powerlaw(n, mean_len, sigma, max_len, seed, _p(ai), _p(aj), _p(av),
                                    C.byref(nnz)))
        return CSR(n, ai, aj, av)

    @staticmethod
    def aniso27(nx, ny=None, nz=None, aniso=100.0, seed=4):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        n = nx * ny * nz
        ai = np.zeros(n + 1, np.int32)
        nnz = C.c_int64()
        check(lib().la_gen_aniso27(nx, ny, nz, aniso, seed, _p(ai), None, None, C.byref(nnz)))
        aj, av = np.zeros(nnz.value, np.int32), np.zeros(nnz.value)
        check(lib().la_gen_aniso27(nx, ny, nz, aniso, seed, _p(ai), _p(aj), _p(av),
                                   C.byref(nnz)))
        return CSR(n, ai, aj, av)

    @staticmethod
    def diffusion2d(nx, ny, theta):
        theta = _f64(theta)
        n = nx * ny
        ne = 2 * nx * ny - nx - ny
        if theta.size != ne:
            raise ValueError("theta must have %d entries" % ne)
        ai = np.zeros(n + 1, np.int32)
        nnz = C.c_int64()
        check(lib().la_gen_diffusion2d(nx, ny, _p(theta), _p(ai), None, None, None, None, None,
                                       C.byref(nnz)))
        aj, av = np.zeros(nnz.value, np.int32), np.zeros(nnz.value)
        ei, ej = np.zeros(ne, np.int32), np.zeros(ne, np.int32)
        vm = np.zeros(4 * ne, np.int32)
        check(lib().la_gen_diffusion2d(nx, ny, _p(theta), _p(ai), _p(aj), _p(av), _p(ei),
                                       _p(ej), _p(vm), C.byref(nnz)))
        return CSR(n, ai, aj, av), ei, ej, vm.reshape(ne, 4)

    @staticmethod
    def uniform(n, seed, lo=-1.0, hi=1.0):
        out = np.zeros(n)
        check(lib().la_gen_uniform(n, seed, lo, hi, _p(out)))
        return out

    @staticmethod
    def lognormal(n, seed, mu=0.0, sigma=1.0):
        out = np.zeros(n)
        check(lib().la_gen_lognormal(n, seed, mu, sigma, _p(out)))
        return out
