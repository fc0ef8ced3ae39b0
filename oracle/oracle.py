"""ctypes front end of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the result to compare against or as the
timed CPU baseline. The
product package (linearalgebra_amd) never imports it.

Every wrapper takes/returns numpy arrays (int32 indices, float64 values) and
keeps the reference's argument meaning: for SpMV `b` is the input vector and
`x` the output, x = alpha*A*b + beta*x (sparse_mat_op/spmv.hpp:10).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = C.CDLL(_SO)
        _declare(_lib)
    return _lib


class _Prec(C.Structure):
    _fields_ = [("kind", C.c_int), ("fai", C.c_void_p), ("faj", C.c_void_p),
                ("fav", C.c_void_p)]


def _declare(L):
    i, d, vp = C.c_int, C.c_double, C.c_void_p
    spmv_args = [i, i, _I, _I, _D, _D, _D, d, d]
    L.orc_spmv_serial.argtypes = spmv_args
    for f in ("orc_spmv_parallel", "orc_spmv_segsum", "orc_spmv_albus"):
        getattr(L, f).argtypes = spmv_args + [i]
    L.orc_forward_sub.argtypes = [i, i, _I, _I, _D, _D, _D]
    L.orc_backward_sub.argtypes = [i, i, _I, _I, _D, _D, _D, _D]
    L.orc_forward_sub_t.argtypes = [i, i, _I, _I, _D, _D, _D]
    L.orc_trsv_jacobi.argtypes = [i, i, _I, _I, _D, C.c_void_p, _D, i, _D]
    L.orc_topo_sort2.argtypes = [i, i, _I, _I, i, _I, _I]
    L.orc_topo_sort2.restype = i
    L.orc_level_forward.argtypes = [_I, _I, i, i, i, _I, _I, _D, _D, _D]
    L.orc_level_backward.argtypes = [_I, _I, i, i, i, _I, _I, _D, _D, _D, _D]
    L.orc_split_ldu.argtypes = [i, i, _I, _I, _D, _I, vp, vp, vp, _I, vp, vp,
                                C.POINTER(i), C.POINTER(i)]
    L.orc_split_triangle.argtypes = [i, i, _I, _I, _D, i, _I, vp, vp,
                                     C.POINTER(i)]
    L.orc_diagonal.argtypes = [i, i, _I, _I, _D, _I, _D]
    L.orc_diagonal.restype = i
    L.orc_transpose.argtypes = [i, i, i, _I, _I, _D, _I, _I, _D]
    L.orc_permute.argtypes = [i, i, _I, _I, _D, vp, vp, _I, _I, _D]
    L.orc_iluk_symbolic.argtypes = [i, i, _I, _I, i, _I, vp, _I]
    L.orc_iluk_symbolic.restype = i
    L.orc_iluk_numeric.argtypes = [i, i, _I, _I, _D, _I, _I, _I, _D]
    L.orc_iluk_numeric.restype = i
    L.orc_icc_symbolic.argtypes = [i, i, _I, _I, _I, i, _I, vp]
    L.orc_icc_symbolic.restype = i
    L.orc_icc_numeric.argtypes = [i, i, _I, _I, _D, _I, d, _I, _I, _D]
    L.orc_icc_numeric.restype = i
    L.orc_ick_numeric.argtypes = [i, i, _I, _I, _D, _I, d, _I, _I, _D]
    L.orc_ick_numeric.restype = i
    L.orc_pcg.argtypes = [i, i, _I, _I, _D, C.POINTER(_Prec), _D, _D, d, d, i,
                          C.POINTER(i), C.POINTER(d)]
    L.orc_pcg.restype = i
    L.orc_fgmres.argtypes = [i, i, _I, _I, _D, C.POINTER(_Prec), _D, _D, i, d,
                             d, i, C.POINTER(i), C.POINTER(d)]
    L.orc_fgmres.restype = i
    L.orc_prec_apply.argtypes = [i, i, C.POINTER(_Prec), _D, _D]
    L.orc_rowwise_sqrt_norm.argtypes = [i, i, _I, _D, _D]
    L.orc_dtad.argtypes = [i, i, _I, _I, _D, _D]
    L.orc_rcm.argtypes = [i, i, _I, _I, _I, _I]
    L.orc_bfs_levels.argtypes = [i, i, _I, _I, i, _I]
    L.orc_bfs_levels.restype = i
    L.orc_wtime.restype = d
    L.orc_max_threads.restype = i
    L.orc_set_plant.argtypes = [i]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class CSR:
    """Plain CSR triple with an explicit index base (matrix_utils.hpp:58-103:
    Base() = ai[0])."""

    def __init__(self, n, ai, aj, av, ncols=None):
        self.n = int(n)
        self.ncols = int(ncols if ncols is not None else n)
        self.ai, self.aj, self.av = _i32(ai), _i32(aj), _f64(av)

    @property
    def base(self):
        return int(self.ai[0])

    @property
    def nnz(self):
        return int(self.ai[self.n] - self.ai[0])

    def to_scipy(self):
        import scipy.sparse as sp
        b = self.base
        return sp.csr_matrix((self.av, self.aj - b, self.ai - b),
                             shape=(self.n, self.ncols))


def spmv(A, b, x=None, alpha=1.0, beta=0.0, alg="serial", nthreads=4):
    b = _f64(b)
    x = np.zeros(A.n) if x is None else _f64(x).copy()
    L = lib()
    if alg == "serial":
        L.orc_spmv_serial(A.n, A.base, A.ai, A.aj, A.av, b, x, alpha, beta)
    else:
        getattr(L, "orc_spmv_" + alg)(A.n, A.base, A.ai, A.aj, A.av, b, x,
                                      alpha, beta, nthreads)
    return x


def forward_sub(L_, b):
    x = np.zeros(L_.n)
    lib().orc_forward_sub(L_.n, L_.base, L_.ai, L_.aj, L_.av, _f64(b), x)
    return x


def backward_sub(U, diag, b):
    x = np.zeros(U.n)
    lib().orc_backward_sub(U.n, U.base, U.ai, U.aj, U.av, _f64(diag), _f64(b),
                           x)
    return x


def forward_sub_t(U, b):
    x = np.zeros(U.n)
    lib().orc_forward_sub_t(U.n, U.base, U.ai, U.aj, U.av, _f64(b), x)
    return x


def trsv_jacobi(T, diag, b, k):
    """k Jacobi sweeps on (D + T) x = b (diag None: unit)."""
    x = np.zeros(T.n)
    d = None if diag is None else _f64(diag)
    lib().orc_trsv_jacobi(T.n, T.base, T.ai, T.aj, T.av,
                          None if d is None else d.ctypes.data_as(C.c_void_p), _f64(b), k, x)
    return x


def topo_sort2(T, upper):
    iperm = np.zeros(max(T.n, 1), np.int32)
    prefix = np.zeros(T.n + 2, np.int32)
    lv = lib().orc_topo_sort2(T.n, T.base, T.ai, T.aj, 1 if upper else 0,
                              iperm, prefix)
    return iperm[:T.n], prefix[:lv + 1], lv


def level_forward(iperm, prefix, lvls, L_, b):
    x = np.zeros(L_.n)
    lib().orc_level_forward(_i32(iperm), _i32(prefix), lvls, L_.n, L_.base,
                            L_.ai, L_.aj, L_.av, _f64(b), x)
    return x


def level_backward(iperm, prefix, lvls, U, diag, b):
    x = np.zeros(U.n)
    lib().orc_level_backward(_i32(iperm), _i32(prefix), lvls, U.n, U.base,
                             U.ai, U.aj, U.av, _f64(diag), _f64(b), x)
    return x


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def split_ldu(A):
    """Returns (L strict lower, D, U strict upper) (matrix_utils.hpp:720)."""
    n, L = A.n, lib()
    Lai, Uai = np.zeros(n + 1, np.int32), np.zeros(n + 1, np.int32)
    ln, un = C.c_int(), C.c_int()
    L.orc_split_ldu(n, A.base, A.ai, A.aj, A.av, Lai, None, None, None, Uai,
                    None, None, C.byref(ln), C.byref(un))
    Laj, Lav = np.zeros(ln.value, np.int32), np.zeros(ln.value)
    Uaj, Uav = np.zeros(un.value, np.int32), np.zeros(un.value)
    D = np.zeros(n)
    L.orc_split_ldu(n, A.base, A.ai, A.aj, A.av, Lai, _ptr(Laj), _ptr(Lav),
                    _ptr(D), Uai, _ptr(Uaj), _ptr(Uav), C.byref(ln),
                    C.byref(un))
    return CSR(n, Lai, Laj, Lav), D, CSR(n, Uai, Uaj, Uav)


def split_triangle(A, upper=True):
    n, L = A.n, lib()
    Tai = np.zeros(n + 1, np.int32)
    t = C.c_int()
    L.orc_split_triangle(n, A.base, A.ai, A.aj, A.av, 1 if upper else 0, Tai,
                         None, None, C.byref(t))
    Taj, Tav = np.zeros(t.value, np.int32), np.zeros(t.value)
    L.orc_split_triangle(n, A.base, A.ai, A.aj, A.av, 1 if upper else 0, Tai,
                         _ptr(Taj), _ptr(Tav), C.byref(t))
    return CSR(n, Tai, Taj, Tav)


def diagonal(A):
    pos = np.zeros(A.n, np.int32)
    d = np.zeros(A.n)
    ok = lib().orc_diagonal(A.n, A.base, A.ai, A.aj, A.av, pos, d)
    return bool(ok), pos, d


def transpose(A):
    ti = np.zeros(A.ncols + 1, np.int32)
    tj = np.zeros(A.nnz, np.int32)
    tv = np.zeros(A.nnz)
    lib().orc_transpose(A.n, A.ncols, A.base, A.ai, A.aj, A.av, ti, tj, tv)
    return CSR(A.ncols, ti, tj, tv, ncols=A.n)


def permute(A, iperm, perm):
    pai = np.zeros(A.n + 1, np.int32)
    paj = np.zeros(A.nnz, np.int32)
    pav = np.zeros(A.nnz)
    ip = None if iperm is None else _i32(iperm)
    pp = None if perm is None else _i32(perm)
    lib().orc_permute(A.n, A.base, A.ai, A.aj, A.av,
                      None if ip is None else _ptr(ip),
                      None if pp is None else _ptr(pp), pai, paj, pav)
    return CSR(A.n, pai, paj, pav)


def iluk_symbolic(A, level):
    n, L = A.n, lib()
    fai = np.zeros(n + 1, np.int32)
    fu = np.zeros(max(n, 1), np.int32)
    nnz = L.orc_iluk_symbolic(n, A.base, A.ai, A.aj, level, fai, None, fu)
    faj = np.zeros(nnz, np.int32)
    L.orc_iluk_symbolic(n, A.base, A.ai, A.aj, level, fai, _ptr(faj), fu)
    return fai, faj, fu[:n]


def iluk(A, level):
    """ILU(k) factor in combined LU storage, same base as A."""
    fai, faj, fu = iluk_symbolic(A, level)
    fav = np.zeros(len(faj))
    ok = lib().orc_iluk_numeric(A.n, A.base, A.ai, A.aj, A.av, fai, faj,
                                _i32(fu), fav)
    return CSR(A.n, fai, faj, fav), fu, bool(ok)


def icc_symbolic(U, diag_pos, lvl):
    n, L = U.n, lib()
    ia = np.zeros(n + 1, np.int32)
    nnz = L.orc_icc_symbolic(n, U.base, U.ai, U.aj, _i32(diag_pos), lvl, ia,
                             None)
    ja = np.zeros(nnz, np.int32)
    L.orc_icc_symbolic(n, U.base, U.ai, U.aj, _i32(diag_pos), lvl, ia,
                       _ptr(ja))
    return ia, ja


def icc(U, lvl, omega=0.0, diag_pos=None):
    """IC(k) of the upper triangle U (incl. diagonal): upper factor R with
    A ~ R^T R, same base as U. diag_pos defaults to U's row starts (the
    diagonal leads each upper row, as tests/precond_test.cpp:72 passes)."""
    dp = U.ai[:-1] if diag_pos is None else diag_pos
    ia, ja = icc_symbolic(U, dp, lvl)
    va = np.zeros(len(ja))
    ok = lib().orc_icc_numeric(U.n, U.base, U.ai, U.aj, U.av, _i32(dp), omega,
                               ia, ja, va)
    return CSR(U.n, ia, ja, va), bool(ok)


def ick(U, lvl, shift=0.0, diag_pos=None):
    """One pass of incomplete_cholesky_k's numeric phase on the IC(lvl)
    pattern of the upper triangle U: (R, ok), additive diagonal shift."""
    dp = U.ai[:-1] if diag_pos is None else diag_pos
    ia, ja = icc_symbolic(U, dp, lvl)
    va = np.zeros(len(ja))
    ok = lib().orc_ick_numeric(U.n, U.base, U.ai, U.aj, U.av, _i32(dp), shift,
                               ia, ja, va)
    return CSR(U.n, ia, ja, va), bool(ok)


def _prec(kind, F):
    P = _Prec()
    P.kind = {"none": 0, "ilu": 1, "ic": 2}[kind]
    if F is not None:
        P.fai, P.faj, P.fav = _ptr(F.ai), _ptr(F.aj), _ptr(F.av)
    return P


def prec_apply(A, kind, F, r):
    z = np.zeros(A.n)
    P = _prec(kind, F)
    lib().orc_prec_apply(A.n, A.base, C.byref(P), _f64(r), z)
    return z


def pcg(A, b, x0=None, kind="none", F=None, rel_tol=1e-8, abs_tol=0.0,
        maxiter=1000):
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    it, rr = C.c_int(), C.c_double()
    P = _prec(kind, F)
    ok = lib().orc_pcg(A.n, A.base, A.ai, A.aj, A.av, C.byref(P), _f64(b), x,
                       rel_tol, abs_tol, maxiter, C.byref(it), C.byref(rr))
    return x, bool(ok), it.value, rr.value


def fgmres(A, b, x0=None, kind="none", F=None, restart=30, rel_tol=1e-8,
           abs_tol=0.0, maxiter=1000):
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    it, rr = C.c_int(), C.c_double()
    P = _prec(kind, F)
    ok = lib().orc_fgmres(A.n, A.base, A.ai, A.aj, A.av, C.byref(P), _f64(b),
                          x, restart, rel_tol, abs_tol, maxiter, C.byref(it),
                          C.byref(rr))
    return x, bool(ok), it.value, rr.value


def rowwise_sqrt_norm(A):
    d = np.zeros(A.n)
    lib().orc_rowwise_sqrt_norm(A.n, A.base, A.ai, A.av, d)
    return d


def dtad(A, d):
    av = A.av.copy()
    lib().orc_dtad(A.n, A.base, A.ai, A.aj, av, _f64(d))
    return CSR(A.n, A.ai, A.aj, av, ncols=A.ncols)


def rcm(A):
    iperm = np.zeros(A.n, np.int32)
    perm = np.zeros(A.n, np.int32)
    lib().orc_rcm(A.n, A.base, A.ai, A.aj, iperm, perm)
    return iperm, perm


def bfs_levels(A, source):
    lv = np.zeros(A.n, np.int32)
    h = lib().orc_bfs_levels(A.n, A.base, A.ai, A.aj, source, lv)
    return h, lv


def wtime():
    return lib().orc_wtime()


# planted mistakes (oracle.h): only the tests that show their checks catch a
# wrong oracle switch one on, always through `planted`
PLANT_SPMV_DROP_LAST = 1
PLANT_RCM_ORIGINAL_DEGREES = 2
PLANT_RCM_WIDTH_RESET = 3


class planted:
    """Context manager: run the oracle with one planted mistake."""

    def __init__(self, which):
        self.which = which

    def __enter__(self):
        lib().orc_set_plant(self.which)
        return self

    def __exit__(self, *exc):
        lib().orc_set_plant(0)
        return False


def max_threads():
    return lib().orc_max_threads()
