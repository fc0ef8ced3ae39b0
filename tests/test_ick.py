"""incomplete_cholesky_k, the additive-shift IC(k) of
mkl_wrapper/incomplete_cholesky.cpp:74-274 (as opposed to the multiplicative
incomplete_cholesky_k_2 at :286-342): the oracle's restatement is pinned
against closed forms, then the product's host factorisation (la_ick_numeric)
must match it bit for bit. No GPU is touched."""
import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import as_lib_csr, from_dense, orc, read_mtx, to_dense


def _grid5(nx, ny, seed, base=0):
    """Symmetric 5-point matrix on an nx x ny grid with random negative
    couplings and a diagonally dominant random diagonal (SPD M-matrix)."""
    rng = np.random.default_rng(seed)
    n = nx * ny
    D = np.zeros((n, n))
    for i in range(n):
        x = i % nx
        if x + 1 < nx:
            D[i, i + 1] = D[i + 1, i] = -rng.uniform(0.2, 1.0)
        if i + nx < n:
            D[i, i + nx] = D[i + nx, i] = -rng.uniform(0.2, 1.0)
    for i in range(n):
        D[i, i] = -D[i].sum() + rng.uniform(0.05, 0.5)
    return D, from_dense(D, base)


def _upper(A):
    return orc.split_triangle(A, True)


def test_oracle_grid_recurrence_bitwise():
    """IC(0) of a 5-point matrix is the textbook recurrence
    d_i = a_ii + s - r_{i-nx,i}^2 - r_{i-1,i}^2, r_ii = sqrt(d_i),
    r_{i,j} = a_ij / r_ii (Meijerink and van der Vorst, 1977). The reference
    visits the rows of column i first-in first-out, so row i-nx (queued when
    its previous column was done) is subtracted before row i-1 (queued when
    row i-1 itself was finished); evaluated in that order, every bit agrees."""
    nx, ny, s = 7, 6, 0.125
    D, A = _grid5(nx, ny, 3)
    n = nx * ny
    R = np.zeros((n, n))
    for i in range(n):
        d = D[i, i] + s
        if i - nx >= 0:
            d = d - R[i - nx, i] * R[i - nx, i]
        if i % nx:
            d = d - R[i - 1, i] * R[i - 1, i]
        R[i, i] = np.sqrt(d)
        if (i % nx) + 1 < nx:
            R[i, i + 1] = D[i, i + 1] / R[i, i]
        if i + nx < n:
            R[i, i + nx] = D[i, i + nx] / R[i, i]
    F, ok = orc.ick(_upper(A), 0, shift=s)
    assert ok
    np.testing.assert_array_equal(to_dense(F), R)
    # the multiplicative variant visits the same rows last-in first-out: the
    # two factors differ in rounding only (tests/precond_test.cpp:118-164)
    G, ok2 = orc.icc(_upper(A), 0)
    F0, _ = orc.ick(_upper(A), 0)
    assert ok2
    np.testing.assert_allclose(F0.av, G.av, rtol=1e-14)
    assert np.any(F0.av != G.av)


@pytest.mark.parametrize("s", [0.0, 0.75])
def test_oracle_tridiagonal_is_cholesky(s):
    """A tridiagonal matrix has no fill: IC(k) of it is the Cholesky factor
    of A + s I for the additive shift s."""
    rng = np.random.default_rng(5)
    n = 40
    off = -rng.uniform(0.1, 1.0, n - 1)
    D = np.diag(off, 1) + np.diag(off, -1) + np.diag(2.2 + rng.uniform(0, 1, n))
    for lvl in (0, 2):
        F, ok = orc.ick(_upper(from_dense(D)), lvl, shift=s)
        assert ok
        L = np.linalg.cholesky(D + s * np.eye(n))
        np.testing.assert_allclose(to_dense(F), L.T, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("s", [0.0, 0.3])
def test_oracle_ic0_reproduces_a_on_pattern(s):
    """IC(0) of an M-matrix: (R^T R)_ij = a_ij + s d_ij wherever R has an
    entry (the dropped fill lies off the pattern)."""
    A = orc.CSR(*[getattr(la.gen.laplace3d(6), f) for f in ("n", "ai", "aj", "av")])
    U = _upper(A)
    F, ok = orc.ick(U, 0, shift=s)
    assert ok
    R = to_dense(F)
    M = R.T @ R
    Ad = to_dense(A) + s * np.eye(A.n)
    mask = (R != 0) | (R.T != 0)
    np.testing.assert_allclose(M[mask], Ad[mask], rtol=1e-13, atol=1e-13)


def test_oracle_zero_pivot_fails_only_here():
    """[[1,1],[1,1]] has a pivot of exactly 0: incomplete_cholesky_k rejects
    a pivot <= 0 (:259) while ICCLevelNumeric rejects only < 0
    (precond.hpp:683)."""
    U = _upper(from_dense(np.array([[1.0, 1.0], [1.0, 1.0]])))
    _, ok = orc.ick(U, 0)
    assert not ok
    _, ok2 = orc.icc(U, 0)
    assert ok2


def _cases():
    return [("poisson2d_20", lambda: orc.CSR(*[getattr(la.gen.poisson2d(20), f)
                                             for f in ("n", "ai", "aj", "av")])),
            ("laplace3d_7", lambda: orc.CSR(*[getattr(la.gen.laplace3d(7), f)
                                            for f in ("n", "ai", "aj", "av")])),
            ("grid5_base1", lambda: _grid5(9, 8, 21, base=1)[1]),
            ("ex5", lambda: read_mtx("ex5.mtx")),
            ("ex5_base1", lambda: read_mtx("ex5.mtx", base=1))]


@pytest.mark.parametrize("s", [0.0, 0.01])
@pytest.mark.parametrize("lvl", [0, 1, 3])
@pytest.mark.parametrize("name,make", _cases(), ids=[c[0] for c in _cases()])
def test_ick_bit_exact(name, make, lvl, s):
    U = _upper(make())
    F1, ok1 = orc.ick(U, lvl, shift=s)
    Ul = as_lib_csr(U)
    dp = Ul.ai[:-1]
    ia, ja = la.icc_symbolic(Ul, dp, lvl)
    F2, ok2 = la.ick_numeric(Ul, dp, ia, ja, s)
    assert ok1 == ok2
    if ok1:
        for f in ("ai", "aj", "av"):
            np.testing.assert_array_equal(getattr(F1, f), getattr(F2, f))


def test_ick_shift_doubles_from_initial():
    """[[1,2],[2,1]]: the second pivot 1 + s - 4/(1 + s) is positive only for
    s > 1, so the retry loop (:259-266) stops at s = 1e-3 * 2^10."""
    U = as_lib_csr(_upper(from_dense(np.array([[1.0, 2.0], [2.0, 1.0]]))))
    R, s = la.ick_shifted(U, 0)
    assert s == 1e-3 * 2 ** 10
    np.testing.assert_allclose(to_dense(R).T @ to_dense(R),
                               np.array([[1.0, 2.0], [2.0, 1.0]]) + s * np.eye(2), rtol=1e-14)
    with pytest.raises(la.LaError):
        la.ick_shifted(U, 0, shift=False)


def test_ick_initial_shift_from_min_diagonal():
    """min a_ii = -1 <= 0 gives s_0 = 1e-3 + 1 (:214-223); the second pivot
    3.001 - 0.25/0.001 is negative, so s doubles once to 2.002, where both
    pivots are positive."""
    D = np.array([[-1.0, 0.5], [0.5, 2.0]])
    U = as_lib_csr(_upper(from_dense(D)))
    R, s = la.ick_shifted(U, 0)
    assert s == 2 * (1e-3 + 1.0)
    F, ok = orc.ick(_upper(from_dense(D)), 0, shift=s)
    assert ok
    np.testing.assert_array_equal(F.av, R.av)


def test_ick_pattern_is_iccs():
    """incomplete_cholesky_k's own symbolic phase gives the ICCLevelSymbolic
    pattern (tests/precond_test.cpp:63-116), so the factor uses it."""
    A = _grid5(6, 5, 9)[1]
    U = _upper(A)
    for lvl in (0, 1, 2, 3):
        F, ok = orc.ick(U, lvl)
        ia, ja = orc.icc_symbolic(U, U.ai[:-1], lvl)
        assert ok
        np.testing.assert_array_equal(F.ai, ia)
        np.testing.assert_array_equal(F.aj, ja)
