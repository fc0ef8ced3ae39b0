"""The 256 x 256 5-point Poisson chain (n = 65,536, nnz = 326,656): SpMV
y = A x with x ~ U[-1, 1), ILU(0) factorisation with a forward/backward
solve, and IC(0)-PCG on b = A 1 to a 1e-8 true relative residual -- the
sequence execs/linearSolve.cpp:82-134 and the reference's tests run.

CPU part: the oracle's chain against scipy (SpMV, the ILU(0) solve via
scipy's triangular solves of the same factors, the PCG solution against a
direct solve, whose exact answer is the ones vector).
GPU part: the HIP chain through the C ABI against the oracle's, FP64,
1e-12 relative for SpMV and the ILU apply, and the PCG solution within the
tolerance-implied error."""
import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import as_lib_csr, orc, rel_err

N = 256


def _chain_inputs():
    A = la.gen.poisson2d(N)
    Ao = orc.CSR(A.n, A.ai, A.aj, A.av)
    x = la.gen.uniform(A.n, 101)
    b = orc.spmv(Ao, np.ones(A.n))
    return A, Ao, x, b


def test_c1_sizes():
    A, _, _, _ = _chain_inputs()
    assert (A.n, A.nnz) == (65536, 326656)


def test_c1_oracle_chain_against_scipy():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from scipy.sparse.linalg import spsolve_triangular
    A, Ao, x, b = _chain_inputs()
    S = sp.csr_matrix((A.av, A.aj, A.ai), shape=(A.n, A.n))
    # SpMV
    assert rel_err(orc.spmv(Ao, x), S @ x) < 1e-15
    # ILU(0): the oracle's apply equals L^-1 then U^-1 of its own factors
    F, _, ok = orc.iluk(Ao, 0)
    assert ok
    Fs = sp.csr_matrix((F.av, F.aj, F.ai), shape=(A.n, A.n))
    L = sp.tril(Fs, -1, format="csr") + sp.identity(A.n, format="csr")
    U = sp.triu(Fs, 0, format="csr")
    z = orc.prec_apply(Ao, "ilu", F, b)
    w = spsolve_triangular(L, b, lower=True)
    zr = spsolve_triangular(U, w, lower=False)
    assert rel_err(z, zr) < 1e-12
    # IC(0)-PCG on b = A 1 to 1e-8: the solution is the ones vector
    R, ok = orc.icc(orc.split_triangle(Ao, True), 0)
    assert ok
    xs, conv, it, rr = orc.pcg(Ao, b, kind="ic", F=R, rel_tol=1e-8, maxiter=2000)
    assert conv and rr < 1e-8
    xd = spl.spsolve(S.tocsc(), b)
    assert rel_err(xd, np.ones(A.n)) < 1e-10
    assert rel_err(xs, xd) < 1e-8 * 1e4  # cond(A) ~ 2.7e4 on this grid


@pytest.mark.gpu
def test_c1_gpu_chain_against_oracle(dev):
    import torch
    A, Ao, x, b = _chain_inputs()
    dA = la.DeviceMatrix(A)
    xt = torch.tensor(x, device=dev)
    y = torch.zeros_like(xt)
    for alg in (la.SPMV_ROW, la.SPMV_ALBUS, la.SPMV_SEGSUM):
        dA.spmv(xt, y, alg=alg)
        assert rel_err(y.cpu().numpy(), orc.spmv(Ao, x)) < 1e-12
    F = orc.iluk(Ao, 0)[0]
    P = la.Preconditioner(as_lib_csr(F), "ilu")
    bt = torch.tensor(b, device=dev)
    z = torch.zeros_like(bt)
    zr = orc.prec_apply(Ao, "ilu", F, b)
    for mode in (la.TRSV_LEVEL, la.TRSV_SYNCFREE):
        P.apply(bt, z, mode=mode)
        assert rel_err(z.cpu().numpy(), zr) < 1e-12
    R, ok = orc.icc(orc.split_triangle(Ao, True), 0)
    Pc = la.Preconditioner(as_lib_csr(R), "ic")
    xo, conv, it_o, _ = orc.pcg(Ao, b, kind="ic", F=R, rel_tol=1e-8, maxiter=2000)
    xg = torch.zeros_like(bt)
    res = la.pcg(dA, Pc, bt, xg, rel_tol=1e-8, max_iters=2000)
    assert res["converged"] and res["rel_residual"] < 1e-8
    assert abs(res["iters"] - it_o) <= 2
    assert rel_err(xg.cpu().numpy(), xo) < 1e-8 * 1e4
