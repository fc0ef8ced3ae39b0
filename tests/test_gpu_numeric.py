"""ILU(k) / IC(k) numeric factorisation on the device (la_prec_create_*_gpu +
la_prec_refactor) against the host factorisation, which restates
mkl_wrapper/incomplete_lu.cpp:277-325 and sparse_mat_op/precond.hpp:615-693
and is itself pinned against the oracle (tests/test_host.py) and brute force
(tests/test_oracle.py). The device factor must be BITWISE the host's (same
operations, same order, no fused multiply-add); the solves built from it
must match the host-factor preconditioner's to 1e-14 and the oracle's apply
to 1e-12."""
import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import as_lib_csr, orc, read_mtx, rel_err

pytestmark = pytest.mark.gpu


def _o(A):
    return orc.CSR(A.n, A.ai, A.aj, A.av, ncols=A.ncols)


def _ilu_cases():
    return [("ex5", lambda: as_lib_csr(read_mtx("ex5.mtx"))),
            ("poisson2d_40", lambda: la.gen.poisson2d(40)),
            ("laplace3d_12", lambda: la.gen.laplace3d(12)),
            ("aniso27_10", lambda: la.gen.aniso27(10, 9, 8)),
            ("powerlaw_3000", lambda: la.gen.powerlaw(3000, 8.0, 1.0, 400, seed=2)),
            ("powerlaw_longrows", lambda: la.gen.powerlaw(4000, 16.0, 2.0, 3000, seed=5))]


def _ic_cases():
    return [("ex5", lambda: as_lib_csr(read_mtx("ex5.mtx"))),
            ("ex5_base1", lambda: as_lib_csr(read_mtx("ex5.mtx", base=1))),
            ("poisson2d_40", lambda: la.gen.poisson2d(40)),
            ("laplace3d_12", lambda: la.gen.laplace3d(12)),
            ("aniso27_10", lambda: la.gen.aniso27(10, 9, 8))]


@pytest.mark.parametrize("name,make", _ilu_cases(), ids=[c[0] for c in _ilu_cases()])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_ilu_gpu_bitwise_host(dev, name, make, level):
    import torch
    A = make()
    if name == "powerlaw_longrows" and level > 0:
        pytest.skip("ILU(k>0) of the long-row matrix fills in beyond a quick test")
    F = la.iluk(A, level)  # host numeric
    P = la.Preconditioner.ilu_gpu(A, level)
    P.refactor(torch.tensor(A.av, device=dev))
    G = P.factor()
    assert np.array_equal(G.ai, np.asarray(F.ai) - F.base)
    assert np.array_equal(G.aj, np.asarray(F.aj) - F.base)
    assert np.array_equal(G.av, F.av), "max diff %.3e" % np.max(np.abs(G.av - F.av))
    # the solves built from the device factor
    r = la.gen.uniform(A.n, 3)
    zr = orc.prec_apply(_o(A), "ilu", _o(F), r)
    rt = torch.tensor(r, device=dev)
    for mode in (la.TRSV_LEVEL, la.TRSV_SYNCFREE):
        z = torch.zeros_like(rt)
        P.apply(rt, z, mode=mode)
        assert rel_err(z.cpu().numpy(), zr) < 1e-12


@pytest.mark.parametrize("name,make", _ic_cases(), ids=[c[0] for c in _ic_cases()])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_ic_gpu_bitwise_host(dev, name, make, level):
    import torch
    A = make()
    U = la.split_triangle(A, upper=True)
    R, ok = la.icc(U, level)
    P = la.Preconditioner.ic_gpu(A, level)
    if not ok:  # ex5 meets a negative pivot in IC(0): the device says so too
        with pytest.raises(la.LaError) as e:
            P.refactor(torch.tensor(A.av, device=dev))
        assert e.value.code == 5
        return
    P.refactor(torch.tensor(A.av, device=dev))
    G = P.factor()
    assert np.array_equal(G.ai, np.asarray(R.ai) - R.base)
    assert np.array_equal(G.aj, np.asarray(R.aj) - R.base)
    assert np.array_equal(G.av, R.av), "max diff %.3e" % np.max(np.abs(G.av - R.av))
    r = la.gen.uniform(A.n, 4)
    zr = orc.prec_apply(_o(A), "ic", _o(R), r)
    rt = torch.tensor(r, device=dev)
    for mode in (la.TRSV_LEVEL, la.TRSV_SYNCFREE):
        z = torch.zeros_like(rt)
        P.apply(rt, z, mode=mode)
        assert rel_err(z.cpu().numpy(), zr) < 1e-12


@pytest.mark.parametrize("name,make", _ic_cases(), ids=[c[0] for c in _ic_cases()])
@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("shift", [0.0, 0.125])
def test_ick_gpu_bitwise_host(dev, name, make, level, shift):
    """incomplete_cholesky_k's numeric phase on the device (additive shift,
    first-in first-out rows; incomplete_cholesky.cpp:195-274) is bitwise the
    host's la_ick_numeric, which is bitwise the oracle's (tests/test_ick.py);
    a pivot <= 0 is reported by both."""
    import torch
    A = make()
    U = la.split_triangle(A, upper=True)
    dp = U.ai[:-1]
    ia, ja = la.icc_symbolic(U, dp, level)
    R, ok = la.ick_numeric(U, dp, ia, ja, shift)
    P = la.Preconditioner.ick_gpu(A, level, shift)
    if not ok:
        with pytest.raises(la.LaError) as e:
            P.refactor(torch.tensor(A.av, device=dev))
        assert e.value.code == 5
        return
    P.refactor(torch.tensor(A.av, device=dev))
    G = P.factor()
    assert np.array_equal(G.ai, np.asarray(R.ai) - R.base)
    assert np.array_equal(G.aj, np.asarray(R.aj) - R.base)
    assert np.array_equal(G.av, R.av), "max diff %.3e" % np.max(np.abs(G.av - R.av))
    r = la.gen.uniform(A.n, 5)
    zr = orc.prec_apply(_o(A), "ic", _o(R), r)
    rt = torch.tensor(r, device=dev)
    z = torch.zeros_like(rt)
    P.apply(rt, z, mode=la.TRSV_SYNCFREE)
    assert rel_err(z.cpu().numpy(), zr) < 1e-12


def test_ick_gpu_zero_pivot(dev):
    """[[1,1],[1,1]]: the second pivot is exactly 0, which
    incomplete_cholesky_k rejects (:259) and ICCLevelNumeric accepts
    (precond.hpp:683)."""
    import torch
    from helpers import from_dense
    A = as_lib_csr(from_dense(np.array([[1.0, 1.0], [1.0, 1.0]])))
    vals = torch.tensor(A.av, device=dev)
    with pytest.raises(la.LaError) as e:
        la.Preconditioner.ick_gpu(A, 0).refactor(vals)
    assert e.value.code == 5
    la.Preconditioner.ic_gpu(A, 0).refactor(vals)


def test_ic_gpu_shift_and_refactor(dev):
    """Two refactors in a row with different values; the multiplicative
    shift (1 + omega) of ICCLevelNumeric (precond.hpp:636-638)."""
    import torch
    A = la.gen.laplace3d(10)
    P = la.Preconditioner.ic_gpu(A, 1, omega=0.05)
    for scale in (1.0, 3.0):
        vals = np.asarray(A.av) * scale
        Ascaled = la.CSR(A.n, A.ai, A.aj, vals)
        R, ok = la.icc(la.split_triangle(Ascaled, upper=True), 1, omega=0.05)
        assert ok
        P.refactor(torch.tensor(vals, device=dev))
        assert np.array_equal(P.factor().av, R.av)


def test_ilu_gpu_refactor_and_transposed_apply(dev):
    """Refactor with new values after a transposed apply built the
    transposed solve plans: those follow the new factor too."""
    import torch
    A = la.gen.powerlaw(2000, 8.0, 1.0, 300, seed=4)
    P = la.Preconditioner.ilu_gpu(A, 1)
    P.refactor(torch.tensor(A.av, device=dev))
    r = torch.tensor(la.gen.uniform(A.n, 8), device=dev)
    z = torch.zeros_like(r)
    P.apply_transpose(r, z)  # builds the transposed plans
    vals2 = np.asarray(A.av) * 1.5  # same pattern, new values
    A2 = la.CSR(A.n, A.ai, A.aj, vals2)
    P.refactor(torch.tensor(vals2, device=dev))
    F2 = la.iluk(A2, 1)
    Ph = la.Preconditioner(F2, "ilu")
    for mode in (la.TRSV_LEVEL, la.TRSV_SYNCFREE):
        z1, z2 = torch.zeros_like(r), torch.zeros_like(r)
        P.apply_transpose(r, z1, mode=mode)
        Ph.apply_transpose(r, z2, mode=mode)
        assert rel_err(z1.cpu().numpy(), z2.cpu().numpy()) < 1e-14


def test_pcg_with_device_factor_matches_host_factor(dev):
    import torch
    A = la.gen.laplace3d(20)
    U = la.split_triangle(A, upper=True)
    R, ok = la.icc(U, 0)
    Ph = la.Preconditioner(R, "ic")
    Pg = la.Preconditioner.ic_gpu(A, 0)
    Pg.refactor(torch.tensor(A.av, device=dev))
    dA = la.DeviceMatrix(A)
    b = torch.tensor(la.gen.uniform(A.n, 9), device=dev)
    res = []
    for P in (Ph, Pg):
        x = torch.zeros_like(b)
        res.append((la.pcg(dA, P, b, x, rel_tol=1e-10, max_iters=1000), x.cpu().numpy()))
    assert res[0][0]["iters"] == res[1][0]["iters"]
    assert np.array_equal(res[0][1], res[1][1])


def test_ic_gpu_negative_pivot(dev):
    import torch
    # symmetric, not positive definite: IC(0) meets a negative pivot
    D = np.array([[1.0, 2.0, 0.0], [2.0, 1.0, 1.0], [0.0, 1.0, 3.0]])
    from helpers import from_dense
    A = as_lib_csr(from_dense(D))
    P = la.Preconditioner.ic_gpu(A, 0)
    with pytest.raises(la.LaError) as e:
        P.refactor(torch.tensor(A.av, device=dev))
    assert e.value.code == 5  # LA_ERR_NUMERIC


def test_ilu_gpu_missing_diagonal_rejected(dev):
    from helpers import from_dense
    A = as_lib_csr(from_dense(np.array([[0.0, 1.0], [1.0, 2.0]])))
    with pytest.raises(la.LaError) as e:
        la.Preconditioner.ilu_gpu(A, 0)
    assert e.value.code == 5
