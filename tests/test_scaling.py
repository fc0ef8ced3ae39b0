"""Symmetric row-norm scaling (mkl_sparse_mat.cpp:517-544): rowwiseSqrtNorm
and DtAD. The oracle restates them (oracle/oracle.c) and is pinned here by
a Python loop in the reference's order (bitwise), by d_i^4 * ||row_i||^2 = 1
within rounding, and by the dense product D A D; the library's host and
device versions must be bitwise the oracle's."""
import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import as_lib_csr, orc, read_mtx


def _cases():
    return [("ex5", read_mtx("ex5.mtx")),
            ("rdist1_base1", read_mtx("rdist1.mtx", base=1)),
            ("aniso27_6", orc.CSR(*_g(la.gen.aniso27(6))))]


def _g(A):
    return A.n, A.ai, A.aj, A.av


def _loop_norm(A):
    d = np.zeros(A.n)
    b = A.base
    for i in range(A.n):
        s = 0.0
        for q in range(A.ai[i] - b, A.ai[i + 1] - b):
            s += A.av[q] * A.av[q]
        d[i] = 1.0 / np.sqrt(np.sqrt(s)) if s > np.finfo(float).tiny else 1.0
    return d


@pytest.mark.parametrize("name,A", _cases(), ids=[c[0] for c in _cases()])
def test_oracle_row_norm_pins(name, A):
    d = orc.rowwise_sqrt_norm(A)
    assert np.array_equal(d, _loop_norm(A))
    b = A.base
    for i in range(A.n):
        s = float(np.sum(A.av[A.ai[i] - b:A.ai[i + 1] - b] ** 2))
        if s > 0:
            assert abs(d[i] ** 4 * s - 1.0) < 1e-14


def test_oracle_dtad_is_dense_product():
    A = read_mtx("ex5.mtx")
    d = orc.rowwise_sqrt_norm(A)
    S = orc.dtad(A, d)
    D = A.to_scipy().toarray()
    ref = np.diag(d) @ D @ np.diag(d)
    got = S.to_scipy().toarray()
    assert np.allclose(got, ref, rtol=1e-15, atol=0)


@pytest.mark.parametrize("name,A", _cases(), ids=[c[0] for c in _cases()])
def test_host_bitwise_oracle(name, A):
    L = as_lib_csr(A)
    d = la.rowwise_sqrt_norm(L)
    assert np.array_equal(d, orc.rowwise_sqrt_norm(A))
    if A.n == A.ncols:
        assert np.array_equal(la.dtad(L, d).av, orc.dtad(A, d).av)


def test_empty_row_gets_one():
    from helpers import from_dense
    A = from_dense(np.array([[2.0, 0.0], [0.0, 0.0]]))
    assert list(la.rowwise_sqrt_norm(as_lib_csr(A))) == [1.0 / np.sqrt(2.0), 1.0]


@pytest.mark.gpu
@pytest.mark.parametrize("name,A", _cases(), ids=[c[0] for c in _cases()])
def test_device_bitwise_host(dev, name, A):
    import torch
    if A.n != A.ncols:
        pytest.skip("DtAD needs a square matrix")
    L = as_lib_csr(A)
    dm = la.DeviceMatrix(L)
    d = torch.zeros(A.n, dtype=torch.float64, device=dev)
    dm.rowwise_sqrt_norm(d)
    dh = la.rowwise_sqrt_norm(L)
    assert np.array_equal(d.cpu().numpy(), dh)
    dm.dtad(d)
    S = la.dtad(L, dh)
    x = torch.tensor(la.gen.uniform(A.n, 2), device=dev)
    y = torch.zeros_like(x)
    dm.spmv(x, y, alg=la.SPMV_ROW)
    ref = orc.spmv(orc.CSR(A.n, S.ai, S.aj, S.av), x.cpu().numpy())
    assert np.allclose(y.cpu().numpy(), ref, rtol=1e-13, atol=1e-300)
