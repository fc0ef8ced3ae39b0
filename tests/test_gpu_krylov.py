"""GPU PCG / FGMRES (mkl_wrapper/mkl_solver.cpp:150-386) against the oracle's
textbook restatements and against closed-form solutions. MKL's RCI
arithmetic is closed, so parity is on the reference's stopping rule: the
final true relative residual ||b - A x|| / ||b|| (FP64) below the tolerance,
and the solution within the tolerance-implied error of the oracle's."""
import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import as_lib_csr, orc, read_mtx

pytestmark = pytest.mark.gpu


def _o(A):
    return orc.CSR(A.n, A.ai, A.aj, A.av, ncols=A.ncols)


def _ic_problem(grid=(24, 24, 24), lvl=0, rcm=False):
    A = _o(la.gen.laplace3d(*grid))
    if rcm:
        ip, p = orc.rcm(A)
        A = orc.permute(A, ip, p)
    U = orc.split_triangle(A, True)
    R, ok = orc.icc(U, lvl)
    assert ok
    return A, R


@pytest.mark.parametrize("mode", [la.TRSV_LEVEL, la.TRSV_SYNCFREE])
@pytest.mark.parametrize("stop", [la.STOP_TRUE_RESIDUAL, la.STOP_RECURRENCE])
def test_pcg_ic0_fp64(dev, mode, stop):
    import torch
    A, R = _ic_problem(rcm=True)
    b = la.gen.uniform(A.n, 21)
    dA = la.DeviceMatrix(as_lib_csr(A))
    P = la.Preconditioner(as_lib_csr(R), "ic")
    x = torch.zeros(A.n, dtype=torch.float64, device=dev)
    res = la.pcg(dA, P, torch.tensor(b, device=dev), x, rel_tol=1e-10, max_iters=2000,
                 trsv_mode=mode, stop_rule=stop)
    print(res, flush=True)
    assert res["converged"] and res["rel_residual"] < 1e-10
    xo, ok, it, rr = orc.pcg(A, b, kind="ic", F=R, rel_tol=1e-10, maxiter=2000)
    assert ok
    # same algorithm, same stopping rule: iteration counts agree closely
    assert abs(res["iters"] - it) <= 2
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-7 * np.linalg.norm(xo)


def test_pcg_ic_k_additive(dev):
    """PCG with the incomplete_cholesky_k factor (additive shift,
    incomplete_cholesky.cpp:195-274) built by the product's host phase on
    the RCM-ordered 24^3 Laplacian; the oracle's PCG with the oracle's
    factor of the same matrix is the reference solution."""
    import torch
    A = _o(la.gen.laplace3d(24, 24, 24))
    ip, p = orc.rcm(A)
    A = orc.permute(A, ip, p)
    U = orc.split_triangle(A, True)
    R, s = la.ick_shifted(as_lib_csr(U), 1)
    Ro, ok = orc.ick(U, 1, shift=s)
    assert ok
    b = la.gen.uniform(A.n, 23)
    dA = la.DeviceMatrix(as_lib_csr(A))
    P = la.Preconditioner(R, "ic")
    x = torch.zeros(A.n, dtype=torch.float64, device=dev)
    res = la.pcg(dA, P, torch.tensor(b, device=dev), x, rel_tol=1e-10, max_iters=2000)
    print(res, flush=True)
    assert res["converged"] and res["rel_residual"] < 1e-10
    xo, ok, it, rr = orc.pcg(A, b, kind="ic", F=Ro, rel_tol=1e-10, maxiter=2000)
    assert ok and abs(res["iters"] - it) <= 2
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-7 * np.linalg.norm(xo)


@pytest.mark.parametrize("prec", [la.PREC_MIXED_F32, la.PREC_MIXED_BF16,
                                  la.PREC_MIXED_F32_BF16P])
def test_pcg_mixed_reaches_fp64_tolerance(dev, prec):
    import torch
    A, R = _ic_problem(rcm=True)
    b = la.gen.uniform(A.n, 21)
    dA = la.DeviceMatrix(as_lib_csr(A))
    P = la.Preconditioner(as_lib_csr(R), "ic")
    x = torch.zeros(A.n, dtype=torch.float64, device=dev)
    res = la.pcg(dA, P, torch.tensor(b, device=dev), x, rel_tol=1e-10, max_iters=4000,
                 precision=prec, inner_tol=1e-4 if prec == la.PREC_MIXED_BF16 else 1e-5)
    print(res, flush=True)
    assert res["converged"] and res["rel_residual"] < 1e-10
    r = orc.spmv(A, x.cpu().numpy(), b, alpha=-1.0, beta=1.0)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10


def test_pcg_max_iters_reports_failure(dev):
    """mkl_solver.cpp:250 returns false when the iteration limit is hit."""
    import torch
    A, R = _ic_problem(grid=(16, 16, 16))
    dA = la.DeviceMatrix(as_lib_csr(A))
    P = la.Preconditioner(as_lib_csr(R), "ic")
    b = torch.tensor(la.gen.uniform(A.n, 1), device=dev)
    x = torch.zeros(A.n, dtype=torch.float64, device=dev)
    res = la.pcg(dA, P, b, x, rel_tol=1e-12, max_iters=3)
    assert not res["converged"] and res["status"] == 4 and res["iters"] == 3
    with pytest.raises(la.LaError):
        la.pcg(dA, P, b, x, rel_tol=1e-12, max_iters=3, raise_on_fail=True)


def test_pcg_zero_rhs(dev):
    import torch
    A, R = _ic_problem(grid=(8, 8, 8))
    dA = la.DeviceMatrix(as_lib_csr(A))
    P = la.Preconditioner(as_lib_csr(R), "ic")
    b = torch.zeros(A.n, dtype=torch.float64, device=dev)
    x = torch.ones(A.n, dtype=torch.float64, device=dev)
    res = la.pcg(dA, P, b, x)
    assert res["converged"] and float(x.abs().max()) == 0.0


def test_cg_distinct_eigenvalues(dev):
    """Unpreconditioned CG on a diagonal SPD matrix with k distinct
    eigenvalues converges in at most k iterations (exact arithmetic)."""
    import torch
    n, k = 4000, 5
    d = np.repeat(np.array([1.0, 2.0, 3.5, 7.0, 11.0]), n // k)
    A = orc.CSR(n, np.arange(n + 1), np.arange(n), d)
    dA = la.DeviceMatrix(as_lib_csr(A))
    b = torch.tensor(la.gen.uniform(n, 2), device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    res = la.pcg(dA, None, b, x, rel_tol=1e-12, max_iters=50)
    assert res["converged"] and res["iters"] <= k + 1


def test_pcg_exact_preconditioner_one_iteration(dev):
    """M = A (tridiagonal SPD: IC(0) is the exact Cholesky factor) -> one
    iteration."""
    import torch
    n = 3000
    main, off = 4.0 * np.ones(n), -1.0 * np.ones(n - 1)
    import scipy.sparse as sp
    S = sp.diags([off, main, off], [-1, 0, 1]).tocsr()
    A = orc.CSR(n, S.indptr, S.indices, S.data)
    U = orc.split_triangle(A, True)
    R, ok = orc.icc(U, 0)
    dA = la.DeviceMatrix(as_lib_csr(A))
    P = la.Preconditioner(as_lib_csr(R), "ic")
    b = torch.tensor(la.gen.uniform(n, 2), device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    res = la.pcg(dA, P, b, x, rel_tol=1e-12, max_iters=10)
    assert res["converged"] and res["iters"] == 1


@pytest.mark.parametrize("mode", [la.TRSV_LEVEL, la.TRSV_SYNCFREE])
def test_fgmres_ilu(dev, mode):
    import torch
    A = _o(la.gen.powerlaw(6000, 10.0, 1.0, 500, seed=12))
    F = orc.iluk(A, 1)[0]
    b = la.gen.uniform(A.n, 4)
    dA = la.DeviceMatrix(as_lib_csr(A))
    P = la.Preconditioner(as_lib_csr(F), "ilu")
    x = torch.zeros(A.n, dtype=torch.float64, device=dev)
    res = la.fgmres(dA, P, torch.tensor(b, device=dev), x, rel_tol=1e-10, restart=30,
                    max_iters=600, trsv_mode=mode)
    print(res, flush=True)
    assert res["converged"] and res["rel_residual"] < 1e-9
    xo, ok, it, rr = orc.fgmres(A, b, kind="ilu", F=F, restart=30, rel_tol=1e-10, maxiter=600)
    assert ok and abs(res["iters"] - it) <= 3
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-8 * np.linalg.norm(xo)


@pytest.mark.parametrize("prec", [la.PREC_MIXED_F32, la.PREC_MIXED_BF16,
                                  la.PREC_MIXED_BF16_F32P, la.PREC_MIXED_F32_BF16P])
def test_fgmres_mixed_ir(dev, prec):
    import torch
    A = _o(la.gen.aniso27(14, 14, 14, aniso=100.0))
    F = orc.iluk(A, 1)[0]
    b = la.gen.uniform(A.n, 4)
    dA = la.DeviceMatrix(as_lib_csr(A))
    P = la.Preconditioner(as_lib_csr(F), "ilu")
    x = torch.zeros(A.n, dtype=torch.float64, device=dev)
    res = la.fgmres(dA, P, torch.tensor(b, device=dev), x, rel_tol=1e-10, restart=30,
                    max_iters=3000, precision=prec, inner_tol=1e-4)
    print(res, flush=True)
    assert res["converged"] and res["rel_residual"] < 1e-10


def test_fgmres_right_operator(dev):
    """mkl_fgmres_solver(A, P, R) (mkl_solver.cpp:319-324): FGMRES on
    A R^-1 u = b returns u; x = R^-1 u solves A x = b (checked against a
    dense solve). R = ILU(0), P = ILU(1)."""
    import torch
    A = la.gen.powerlaw(1500, 8.0, 1.0, 200, seed=8)
    R = la.Preconditioner(la.iluk(A, 0), "ilu")
    P = la.Preconditioner(la.iluk(A, 1), "ilu")
    b = torch.tensor(la.gen.uniform(A.n, 12), device=dev)
    u = torch.zeros_like(b)
    res = la.fgmres_right(la.DeviceMatrix(A), P, R, b, u, rel_tol=1e-12, restart=30,
                          max_iters=500)
    assert res["converged"] and res["rel_residual"] < 1e-12
    x = torch.zeros_like(b)
    R.apply(u, x)
    D = orc.CSR(A.n, A.ai, A.aj, A.av).to_scipy().toarray()
    xr = np.linalg.solve(D, b.cpu().numpy())
    assert np.linalg.norm(x.cpu().numpy() - xr) <= 1e-9 * np.linalg.norm(xr)
    with pytest.raises(la.LaError):
        la.fgmres_right(la.DeviceMatrix(A), P, R, b, u, precision=la.PREC_MIXED_F32)


def test_unknown_precision_rejected(dev):
    import torch
    A = la.gen.poisson2d(8)
    dA = la.DeviceMatrix(A)
    b = torch.ones(A.n, dtype=torch.float64, device=dev)
    x = torch.zeros_like(b)
    with pytest.raises(la.LaError) as e:
        la.pcg(dA, None, b, x, raise_on_fail=True, precision=9)
    assert e.value.code == 1  # LA_ERR_ARG


def test_pcg_ex5_reference_fixture(dev):
    """ex5 with IC(3) as the reference's precond tests build it
    (tests/precond_test.cpp:66-75); the reference's default rel_tol 1e-8
    (mkl_solver.h:53)."""
    import torch
    A = read_mtx("ex5.mtx")
    U = orc.split_triangle(A, True)
    R, ok = orc.icc(U, 3)
    dA = la.DeviceMatrix(as_lib_csr(A))
    P = la.Preconditioner(as_lib_csr(R), "ic")
    b = torch.ones(A.n, dtype=torch.float64, device=dev)
    x = torch.zeros(A.n, dtype=torch.float64, device=dev)
    res = la.pcg(dA, P, b, x, rel_tol=1e-8, max_iters=100)
    xo, ok2, it, rr = orc.pcg(A, np.ones(A.n), kind="ic", F=R, rel_tol=1e-8, maxiter=100)
    assert res["converged"] and ok2 and res["iters"] == it


@pytest.mark.parametrize("restart", [0, 100])
def test_fgmres_long_restart(dev, restart):
    """GMRES restart lengths above 63 and restart 0, which the reference
    turns into unrestarted GMRES up to the iteration limit
    (mkl_solver.cpp:267-269, 293-294): unpreconditioned 2-D Poisson 48^2
    needs well over 64 Krylov vectors; the product's iteration count
    follows the oracle's within 3 and both solutions agree."""
    import torch
    A = _o(la.gen.poisson2d(48))
    b = la.gen.uniform(A.n, 31)
    dA = la.DeviceMatrix(as_lib_csr(A))
    x = torch.zeros(A.n, dtype=torch.float64, device=dev)
    res = la.fgmres(dA, None, torch.tensor(b, device=dev), x, rel_tol=1e-10, max_iters=400,
                    restart=restart)
    xo, ok, it, rr = orc.fgmres(A, b, restart=restart, rel_tol=1e-10, maxiter=400)
    print(res, "oracle", ok, it, rr, flush=True)
    assert ok and it > 64
    assert res["converged"] and res["rel_residual"] < 1.01e-10
    assert abs(res["iters"] - it) <= 3
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-6 * np.linalg.norm(xo)
