"""TRSV_JACOBI (k Jacobi sweeps per triangle; no counterpart in the
reference) on the device against the oracle's orc_trsv_jacobi, whose pins
are in test_jacobi_oracle.py. FP64: norm-wise relative error <= 1e-12 (the
sweeps are SpMVs; only FMA contraction and summation order may differ);
k >= levels - 1 must reproduce the exact substitution to the same bound.
FP32 / BF16 factors: 1e-5 / 3e-2 relative against the FP64 oracle."""
import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import as_lib_csr, orc, read_mtx, rel_err

pytestmark = pytest.mark.gpu


def _o(A):
    return orc.CSR(A.n, A.ai, A.aj, A.av, ncols=A.ncols)


def _factors():
    return [("ex5_ilu5", lambda: orc.iluk(read_mtx("ex5.mtx"), 5)[0]),
            ("ex5_ilu5_base1", lambda: orc.iluk(read_mtx("ex5.mtx", base=1), 5)[0]),
            ("poisson2d_48_ilu0", lambda: orc.iluk(_o(la.gen.poisson2d(48)), 0)[0]),
            ("laplace3d_14_ilu1", lambda: orc.iluk(_o(la.gen.laplace3d(14)), 1)[0]),
            ("aniso27_8_ilu1", lambda: orc.iluk(_o(la.gen.aniso27(8)), 1)[0]),
            ("powerlaw_ilu0",
             lambda: orc.iluk(_o(la.gen.powerlaw(3000, 8.0, 1.0, 400, seed=2)), 0)[0])]


@pytest.mark.parametrize("name,make", _factors(), ids=[c[0] for c in _factors()])
@pytest.mark.parametrize("k", [0, 1, 3])
def test_sweeps_match_oracle(dev, name, make, k):
    import torch
    F = make()
    L, D, U = orc.split_ldu(F)
    print("case", name, "k", k, "n", F.n, flush=True)
    b = la.gen.uniform(F.n, 3)
    bt = torch.tensor(b, device=dev)
    fw = la.TriangularSolve(as_lib_csr(L), None, upper=False).set_sweeps(k)
    bw = la.TriangularSolve(as_lib_csr(U), D, upper=True).set_sweeps(k)
    for T, ref in ((fw, orc.trsv_jacobi(L, None, b, k)), (bw, orc.trsv_jacobi(U, D, b, k))):
        x = torch.zeros(F.n, dtype=torch.float64, device=dev)
        T.solve(bt, x, mode=la.TRSV_JACOBI)
        assert rel_err(x.cpu().numpy(), ref) <= 1e-12, name


@pytest.mark.parametrize("name,make", _factors()[:3], ids=[c[0] for c in _factors()[:3]])
def test_enough_sweeps_is_exact(dev, name, make):
    import torch
    F = make()
    L, D, U = orc.split_ldu(F)
    b = la.gen.uniform(F.n, 4)
    bt = torch.tensor(b, device=dev)
    for T, ref in ((la.TriangularSolve(as_lib_csr(L), None, upper=False),
                    orc.forward_sub(L, b)),
                   (la.TriangularSolve(as_lib_csr(U), D, upper=True),
                    orc.backward_sub(U, D, b))):
        T.set_sweeps(T.info()["levels"] - 1)
        x = torch.zeros(F.n, dtype=torch.float64, device=dev)
        T.solve(bt, x, mode=la.TRSV_JACOBI)
        assert rel_err(x.cpu().numpy(), ref) <= 1e-12, name


@pytest.mark.parametrize("k", [1, 3])
def test_ic_apply_and_transpose(dev, k):
    """IC(0) apply M^-1 = R^-1 R^-T with k sweeps per triangle equals the
    oracle's sweeps on R^T then R; M^-T is the same operator."""
    import torch
    A = _o(la.gen.laplace3d(12))
    R, ok = orc.icc(orc.split_triangle(A, True), 0)
    assert ok
    Rl, D, Ru = orc.split_ldu(R)  # R upper: strict part Ru, diagonal D
    Rt = orc.transpose(Ru)        # strict lower of R^T
    b = la.gen.uniform(A.n, 6)
    ref = orc.trsv_jacobi(Ru, D, orc.trsv_jacobi(Rt, D, b, k), k)
    P = la.Preconditioner(as_lib_csr(R), "ic").set_sweeps(k)
    bt = torch.tensor(b, device=dev)
    z = torch.zeros_like(bt)
    P.apply(bt, z, mode=la.TRSV_JACOBI)
    assert rel_err(z.cpu().numpy(), ref) <= 1e-12
    zt = torch.zeros_like(bt)
    P.apply_transpose(bt, zt, mode=la.TRSV_JACOBI)
    assert rel_err(zt.cpu().numpy(), ref) <= 1e-12
    # symmetric operator: <M^-1 u, v> == <u, M^-1 v>
    u, v = la.gen.uniform(A.n, 7), la.gen.uniform(A.n, 8)
    Mu, Mv = torch.zeros_like(bt), torch.zeros_like(bt)
    P.apply(torch.tensor(u, device=dev), Mu, mode=la.TRSV_JACOBI)
    P.apply(torch.tensor(v, device=dev), Mv, mode=la.TRSV_JACOBI)
    a1, a2 = float(Mu.cpu().numpy() @ v), float(u @ Mv.cpu().numpy())
    assert abs(a1 - a2) <= 1e-12 * max(abs(a1), 1.0)


def test_ilu_apply_transpose_sweeps(dev):
    """ILU(1) M^-T r with sweeps: (L U)^T z = r as U^T then L^T, each by
    k sweeps; against the oracle's sweeps on the transposed triangles."""
    import torch
    F = orc.iluk(_o(la.gen.powerlaw(2000, 8.0, 1.0, 300, seed=4)), 1)[0]
    L, D, U = orc.split_ldu(F)
    k = 2
    b = la.gen.uniform(F.n, 8)
    t = orc.trsv_jacobi(orc.transpose(U), D, b, k)        # U^T (lower, diag D)
    ref = orc.trsv_jacobi(orc.transpose(L), None, t, k)   # L^T (upper, unit)
    P = la.Preconditioner(as_lib_csr(F), "ilu").set_sweeps(k)
    z = torch.zeros(F.n, dtype=torch.float64, device=dev)
    P.apply_transpose(torch.tensor(b, device=dev), z, mode=la.TRSV_JACOBI)
    assert rel_err(z.cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("val,tol", [(la.F32, 1e-5), (la.BF16, 3e-2)])
def test_lowprec_sweeps(dev, val, tol):
    import torch
    A = _o(la.gen.laplace3d(16))
    R, ok = orc.icc(orc.split_triangle(A, True), 0)
    Rl, D, Ru = orc.split_ldu(R)
    b = la.gen.uniform(A.n, 6)
    ref = orc.trsv_jacobi(Ru, D, orc.trsv_jacobi(orc.transpose(Ru), D, b, 3), 3)
    P = la.Preconditioner(as_lib_csr(R), "ic")
    for vdt in (torch.float32, torch.float64):
        z = torch.zeros(A.n, dtype=vdt, device=dev)
        P.apply(torch.tensor(b, device=dev, dtype=vdt), z, mode=la.TRSV_JACOBI, val=val)
        assert rel_err(z.double().cpu().numpy(), ref) < tol, (val, vdt)


def test_device_factor_sweeps_follow_refactor(dev):
    """Sweeps with a factor computed on the device follow each refactor,
    and equal the sweeps with the same factor computed on the host."""
    import torch
    A = la.gen.laplace3d(10)
    Pg = la.Preconditioner.ic_gpu(A, 1)
    r = torch.tensor(la.gen.uniform(A.n, 2), device=dev)
    for scale in (1.0, 2.0):
        vals = np.asarray(A.av) * scale
        Pg.refactor(torch.tensor(vals, device=dev))
        R, ok = la.icc(la.split_triangle(la.CSR(A.n, A.ai, A.aj, vals), upper=True), 1)
        Ph = la.Preconditioner(R, "ic")
        z1, z2 = torch.zeros_like(r), torch.zeros_like(r)
        Pg.apply(r, z1, mode=la.TRSV_JACOBI)
        Ph.apply(r, z2, mode=la.TRSV_JACOBI)
        assert rel_err(z1.cpu().numpy(), z2.cpu().numpy()) < 1e-14, scale


@pytest.mark.parametrize("precision", [la.PREC_FP64, la.PREC_MIXED_F32, la.PREC_MIXED_BF16])
def test_pcg_with_sweeps_reaches_tolerance(dev, precision):
    """IC(0)-PCG with 3-sweep triangular solves reaches a 1e-10 true FP64
    residual, in FP64 and in mixed precision with FP64 refinement; the
    operator stays SPD, so the iteration count stays near the exact-solve
    one."""
    import torch
    A = la.gen.laplace3d(24)
    R, ok = la.icc(la.split_triangle(A, upper=True), 0)
    P = la.Preconditioner(R, "ic")
    dA = la.DeviceMatrix(A)
    b = torch.tensor(la.gen.uniform(A.n, 9), device=dev)
    x = torch.zeros_like(b)
    ex = la.pcg(dA, P, b, x, rel_tol=1e-10, max_iters=2000)
    x.zero_()
    res = la.pcg(dA, P, b, x, rel_tol=1e-10, max_iters=2000, trsv_mode=la.TRSV_JACOBI,
                 precision=precision, inner_tol=1e-4)
    assert res["converged"] and res["rel_residual"] <= 1e-10
    r = orc.spmv(_o(A), x.cpu().numpy())
    bh = b.cpu().numpy()
    assert np.linalg.norm(bh - r) <= 1.01e-10 * np.linalg.norm(bh)
    assert res["iters"] <= 2 * ex["iters"]


@pytest.mark.parametrize("k", [0, 1, 4])
def test_tiny_and_long_row_triangles(dev, k):
    """1x1 triangles (no entries: diagonal scaling or a copy), and a
    triangle with a row longer than 64 entries, which has no sliced-ELL
    form (the last sweep then scales in a separate pass)."""
    import torch
    one = orc.CSR(1, np.array([0, 0], np.int32), np.zeros(0, np.int32), np.zeros(0))
    b1 = np.array([3.0])
    x = torch.zeros(1, dtype=torch.float64, device=dev)
    T = la.TriangularSolve(as_lib_csr(one), np.array([2.0]), upper=True).set_sweeps(k)
    T.solve(torch.tensor(b1, device=dev), x, mode=la.TRSV_JACOBI)
    assert x.item() == 1.5
    T = la.TriangularSolve(as_lib_csr(one), None, upper=False).set_sweeps(k)
    T.solve(torch.tensor(b1, device=dev), x, mode=la.TRSV_JACOBI)
    assert x.item() == 3.0
    # strict lower triangle: row n-1 holds columns 0..99 (> 64 entries)
    n = 300
    rng = np.random.default_rng(11)
    rows = [np.arange(max(0, i - 2), i) for i in range(n - 1)] + [np.arange(100)]
    ai = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    aj = np.concatenate(rows).astype(np.int32)
    L = orc.CSR(n, ai, aj, rng.uniform(-0.3, 0.3, len(aj)))
    d = rng.uniform(1.0, 2.0, n)
    b = la.gen.uniform(n, 12)
    Tl = la.TriangularSolve(as_lib_csr(L), d, upper=False).set_sweeps(k)
    xt = torch.zeros(n, dtype=torch.float64, device=dev)
    Tl.solve(torch.tensor(b, device=dev), xt, mode=la.TRSV_JACOBI)
    assert rel_err(xt.cpu().numpy(), orc.trsv_jacobi(L, d, b, k)) <= 1e-12
