"""Pins of the oracle's Jacobi-sweep triangular solve (orc_trsv_jacobi; the
device's TRSV_JACOBI mode, which has no counterpart in the reference):

* D^-1 T is nilpotent with index = number of levels, so k >= levels - 1
  sweeps reproduce the exact substitutions orc_forward_sub /
  orc_backward_sub (themselves pinned against scipy and the reference's
  fixtures);
* k = 0 is D^-1 b;
* for small k, the truncated Neumann series sum_{j<=k} (-D^-1 T)^j D^-1 b
  formed densely with numpy matrix powers.
"""
import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import orc, read_mtx


def _o(A):
    return orc.CSR(A.n, A.ai, A.aj, A.av, ncols=A.ncols)


def _factors():
    return [("ex5_ilu5", lambda: orc.iluk(read_mtx("ex5.mtx"), 5)[0]),
            ("ex5_ilu5_base1", lambda: orc.iluk(read_mtx("ex5.mtx", base=1), 5)[0]),
            ("poisson2d_20_ilu0", lambda: orc.iluk(_o(la.gen.poisson2d(20)), 0)[0]),
            ("aniso27_6_ilu1", lambda: orc.iluk(_o(la.gen.aniso27(6)), 1)[0])]


@pytest.mark.parametrize("name,make", _factors(), ids=[c[0] for c in _factors()])
def test_enough_sweeps_is_exact(name, make):
    F = make()
    L, D, U = orc.split_ldu(F)
    b = la.gen.uniform(F.n, 3)
    lv_l = orc.topo_sort2(L, False)[2]
    lv_u = orc.topo_sort2(U, True)[2]
    xl = orc.trsv_jacobi(L, None, b, lv_l - 1)
    xu = orc.trsv_jacobi(U, D, b, lv_u - 1)
    rl, ru = orc.forward_sub(L, b), orc.backward_sub(U, D, b)
    assert np.linalg.norm(xl - rl) <= 1e-12 * np.linalg.norm(rl)
    assert np.linalg.norm(xu - ru) <= 1e-12 * np.linalg.norm(ru)


@pytest.mark.parametrize("k", [0, 1, 7, 19, 25])
def test_bidiagonal_closed_form(k):
    """L = I - (subdiagonal of ones), b = e_0: the exact solution is all
    ones, and k sweeps reach exactly the first k+1 entries (a chain of 20
    levels needs 19 sweeps). The same for the upper orientation with a
    diagonal of 2: x_i = 2^-(n-i) after enough sweeps."""
    n = 20
    ai = np.concatenate([[0], np.arange(n)]).astype(np.int32)  # row i>0: one entry
    aj = np.arange(n - 1, dtype=np.int32)
    L = orc.CSR(n, ai, aj, -np.ones(n - 1))
    b = np.zeros(n)
    b[0] = 1.0
    x = orc.trsv_jacobi(L, None, b, k)
    ref = (np.arange(n) <= k).astype(float)
    np.testing.assert_array_equal(x, ref)
    # upper: row i has column i+1 with -1, diagonal 2, b = e_{n-1}
    ui = np.concatenate([np.arange(n), [n - 1]]).astype(np.int32)
    uj = np.arange(1, n, dtype=np.int32)
    U = orc.CSR(n, ui, uj, -np.ones(n - 1))
    bu = np.zeros(n)
    bu[-1] = 1.0
    xu = orc.trsv_jacobi(U, np.full(n, 2.0), bu, k)
    d = n - 1 - np.arange(n)  # distance from the last row
    refu = np.where(d <= k, 0.5 ** (d + 1), 0.0)
    np.testing.assert_array_equal(xu, refu)


@pytest.mark.parametrize("k", [0, 1, 2, 5])
def test_neumann_series_dense(k):
    F = orc.iluk(_o(la.gen.poisson2d(9)), 0)[0]
    L, D, U = orc.split_ldu(F)
    b = la.gen.uniform(F.n, 4)
    Ud = U.to_scipy().toarray()
    M = -(Ud / D[:, None])  # -D^-1 U
    y0 = b / D
    ref = sum(np.linalg.matrix_power(M, j) @ y0 for j in range(k + 1))
    got = orc.trsv_jacobi(U, D, b, k)
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-14 * np.abs(ref).max())
    Ld = L.to_scipy().toarray()
    ref_l = sum(np.linalg.matrix_power(-Ld, j) @ b for j in range(k + 1))
    assert np.allclose(orc.trsv_jacobi(L, None, b, k), ref_l, rtol=1e-13,
                       atol=1e-14 * np.abs(ref_l).max())


def test_zero_sweeps_is_diagonal_scaling():
    F = orc.iluk(read_mtx("ex5.mtx"), 1)[0]
    L, D, U = orc.split_ldu(F)
    b = la.gen.uniform(F.n, 5)
    np.testing.assert_array_equal(orc.trsv_jacobi(U, D, b, 0), (1.0 / D) * b)
    np.testing.assert_array_equal(orc.trsv_jacobi(L, None, b, 0), b)
