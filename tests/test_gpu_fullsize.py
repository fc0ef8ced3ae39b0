"""Parity at the full benchmark sizes, through the C ABI.

* 128^3 7-point Laplacian (RCM), n = 2,097,152: SpMV in every form against
  the oracle's ParallelSPMV (bitwise equal to SerialSPMV), FP64 1e-12
  element-wise of (|A||x|)_i; the IC(0) apply (both modes) against the
  oracle's serial apply, 1e-12 relative; IC(0)-PCG to 1e-10 against the
  oracle's run stored in tests/golden/c2_ic0_pcg.json (177 iterations):
  iteration count within 2, true residual < 1e-10, x within 1e-8 relative
  (both runs stop at ~9e-11 on the same Krylov path; cond(A) ~ 6.6e3).
* Power-law matrix, n = 4,000,000, nnz ~ 64.0M: SpMV FP64 (1e-12 as
  above), FP32 values (2e-6 of (|A||x|)_i) and BF16 values (2^-8 + 2e-6),
  FP32 accumulation, against the FP64 oracle; the low-precision forms also
  against the oracle on the rounded values (2e-6: FP32 accumulation only).
* 27-point anisotropic 160^3, n = 4,096,000, nnz = 109,215,352: SpMV FP64
  (FP32/BF16 values at the storage bounds above) and the ILU(1) apply (both
  modes) against the oracle, 1e-12.
* The SpMV forms include the sliced-ELL kernel with 16-bit column offsets
  that AUTO (and the bench line) runs on both grids; 128^3 also checks the
  3-sweep Jacobi IC(0) apply against the oracle's sweeps (1e-12) and mixed
  BF16 PCG with it to a 1e-10 FP64 true residual, recomputed on the host.
"""
import json
import os

import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import GOLDEN, as_lib_csr, bf16_rne, orc, rel_err

BF16_BOUND = 2.0 ** -8 + 2e-6

pytestmark = pytest.mark.gpu


def _o(A):
    return orc.CSR(A.n, A.ai, A.aj, A.av, ncols=A.ncols)


def _absbound(A, x):
    return orc.spmv(orc.CSR(A.n, A.ai, A.aj, np.abs(A.av), ncols=A.ncols), np.abs(x),
                    alg="parallel", nthreads=orc.max_threads())


def _check_spmv(dev, A, algs, dtypes, seed=11):
    import torch
    d = la.DeviceMatrix(A)
    x = la.gen.uniform(A.n, seed)
    ref = orc.spmv(_o(A), x, alg="parallel", nthreads=orc.max_threads())
    bound = _absbound(A, x)
    for dt, tol in dtypes:
        if dt != la.F64:
            d.add_dtype(dt)
            av = A.av.astype(np.float32).astype(np.float64) if dt == la.F32 else bf16_rne(A.av)
            ref_r = orc.spmv(orc.CSR(A.n, A.ai, A.aj, av, ncols=A.ncols), x, alg="parallel",
                             nthreads=orc.max_threads())
        vec = torch.float64 if dt == la.F64 else torch.float32
        xt = torch.tensor(x, device=dev).to(vec)
        for alg in algs:
            y = torch.zeros(A.n, dtype=vec, device=dev)
            d.spmv(xt, y, alg=alg, val=dt, acc=la.F64 if dt == la.F64 else la.F32)
            yh = y.double().cpu().numpy()
            print("n=%d alg=%d dtype=%d max err/bound=%.3e" % (
                A.n, alg, dt, np.max(np.abs(yh - ref) / np.maximum(bound, 1e-300))), flush=True)
            assert np.all(np.abs(yh - ref) <= tol * bound + 1e-300), (alg, dt)
            if dt != la.F64:  # the same values rounded by the tests: accumulation only
                assert np.all(np.abs(yh - ref_r) <= 2e-6 * bound + 1e-300), (alg, dt)
    return d


def _c2():
    A = la.gen.laplace3d(128)
    ip, p = la.rcm(A)
    return la.permute(A, ip, p)


def test_c2_spmv_full(dev):
    A = _c2()
    d = _check_spmv(dev, A, [la.SPMV_ROW, la.SPMV_SEGSUM, la.SPMV_ALBUS, la.SPMV_SELL,
                             la.SPMV_AUTO],
                    [(la.F64, 1e-12), (la.F32, 2e-6), (la.BF16, BF16_BOUND)])
    # the bench line's kernel: sliced ELL with 16-bit column offsets
    assert d.resolve(la.SPMV_AUTO) == la.SPMV_SELL and d.sell_info()["col16"]


def test_c2_ic0_apply_and_pcg_full(dev):
    import torch
    A = _c2()
    U = la.split_triangle(A, upper=True)
    R, ok = la.icc(U, 0)
    assert ok
    P = la.Preconditioner(R, "ic")
    r = la.gen.uniform(A.n, 5)
    zr = orc.prec_apply(_o(A), "ic", _o(R), r)
    rt = torch.tensor(r, device=dev)
    for mode in (la.TRSV_LEVEL, la.TRSV_SYNCFREE):
        z = torch.zeros_like(rt)
        P.apply(rt, z, mode=mode)
        err = rel_err(z.cpu().numpy(), zr)
        print("IC(0) apply mode=%d rel err %.3e" % (mode, err), flush=True)
        assert err < 1e-12
    gold = json.load(open(os.path.join(GOLDEN, "c2_ic0_pcg.json")))
    b = torch.tensor(la.gen.uniform(A.n, 2024), device=dev)
    x = torch.zeros_like(b)
    res = la.pcg(la.DeviceMatrix(A), P, b, x, rel_tol=1e-10, max_iters=5000)
    print(res, "oracle iters", gold["iters"], flush=True)
    assert res["converged"] and res["rel_residual"] < 1e-10
    assert abs(res["iters"] - gold["iters"]) <= 2
    xh = x.cpu().numpy()
    assert abs(np.linalg.norm(xh) - gold["x_norm"]) <= 1e-8 * gold["x_norm"]
    xs = np.array(gold["sample_x"])
    assert np.max(np.abs(xh[gold["sample_index"]] - xs)) <= 1e-8 * np.max(np.abs(xs))


def test_c3_powerlaw_spmv_full(dev):
    A = la.gen.powerlaw(4_000_000, 16.0, 1.0, 20000, seed=3)
    assert A.n == 4_000_000 and 60e6 < A.nnz < 70e6
    d = _check_spmv(dev, A, [la.SPMV_AUTO, la.SPMV_ALBUS, la.SPMV_ROW],
                    [(la.F64, 1e-12), (la.F32, 2e-6), (la.BF16, BF16_BOUND)])
    assert d.resolve(la.SPMV_AUTO) == la.SPMV_COLBLOCK


def test_c4_spmv_and_ilu1_apply_full(dev):
    import torch
    A = la.gen.aniso27(160)
    assert (A.n, A.nnz) == (4_096_000, 109_215_352)
    d = _check_spmv(dev, A, [la.SPMV_ALBUS, la.SPMV_AUTO],
                    [(la.F64, 1e-12), (la.F32, 2e-6), (la.BF16, BF16_BOUND)])
    assert d.resolve(la.SPMV_AUTO) == la.SPMV_SELL and d.sell_info()["col16"]
    del d
    F = la.iluk(A, 1)
    P = la.Preconditioner(F, "ilu")
    print("ILU(1) tiles", P.tiles(), flush=True)
    assert min(P.tiles()) > 0  # the box-grid tiled schedule
    r = la.gen.uniform(A.n, 6)
    zr = orc.prec_apply(_o(A), "ilu", _o(F), r)
    rt = torch.tensor(r, device=dev)
    for mode in (la.TRSV_LEVEL, la.TRSV_SYNCFREE):
        z = torch.zeros_like(rt)
        P.apply(rt, z, mode=mode)
        err = rel_err(z.cpu().numpy(), zr)
        print("ILU(1) apply mode=%d rel err %.3e" % (mode, err), flush=True)
        assert err < 1e-12


def test_c2_jacobi_apply_and_mixed_pcg_full(dev):
    """128^3: the 3-sweep IC(0) apply against the oracle's sweeps (1e-12),
    and mixed BF16 PCG with those sweeps (reliable updates) to a 1e-10 FP64
    true residual, checked again on the host with the oracle's SpMV."""
    import torch
    A = _c2()
    R, ok = la.icc(la.split_triangle(A, upper=True), 0)
    P = la.Preconditioner(R, "ic").set_sweeps(3)
    Ro = _o(R)
    Rl, D, Ru = orc.split_ldu(Ro)
    r = la.gen.uniform(A.n, 8)
    ref = orc.trsv_jacobi(Ru, D, orc.trsv_jacobi(orc.transpose(Ru), D, r, 3), 3)
    rt = torch.tensor(r, device=dev)
    z = torch.zeros_like(rt)
    P.apply(rt, z, mode=la.TRSV_JACOBI)
    err = rel_err(z.cpu().numpy(), ref)
    print("Jacobi-3 IC(0) apply rel err %.3e" % err, flush=True)
    assert err < 1e-12
    b = torch.tensor(la.gen.uniform(A.n, 2024), device=dev)
    x = torch.zeros_like(b)
    res = la.pcg(la.DeviceMatrix(A), P, b, x, rel_tol=1e-10, max_iters=5000,
                 precision=la.PREC_MIXED_BF16, trsv_mode=la.TRSV_JACOBI)
    print(res, flush=True)
    assert res["converged"] and res["rel_residual"] < 1e-10
    bh = b.cpu().numpy()
    rr = bh - orc.spmv(_o(A), x.cpu().numpy(), alg="parallel", nthreads=orc.max_threads())
    assert np.linalg.norm(rr) <= 1.01e-10 * np.linalg.norm(bh)


def test_c4_fgmres_full(dev):
    """The 160^3 27-point solve (mkl_solver.cpp:253-386 with
    incomplete_lu_k at level 1): ILU(1)-FGMRES(30) to a 1e-10 relative
    residual in plain FP64, and mixed precision (FP32 matrix + FP32
    factors; BF16 matrix + FP32 factors) with FP64 iterative refinement,
    each with the exact sync-free triangular solves and with 6 Jacobi
    sweeps per triangle. Every final true residual is recomputed on the
    host with the oracle's ParallelSPMV (<= 1.01e-10 ||b||); every mixed
    solution is compared with the FP64 solution with the same triangular
    solves (both stop at a 1e-10 residual of this anisotropic operator)."""
    import torch
    A = la.gen.aniso27(160)
    P = la.Preconditioner.ilu_gpu(A, 1)
    P.refactor(torch.tensor(A.av, device=dev))
    P.set_sweeps(6)
    dA = la.DeviceMatrix(A)
    bh = la.gen.uniform(A.n, 4)
    b = torch.tensor(bh, device=dev)
    Ao = _o(A)
    J = dict(trsv_mode=la.TRSV_JACOBI)
    S = dict(trsv_mode=la.TRSV_SYNCFREE)
    variants = [("fp64", dict(precision=la.PREC_FP64, **S)),
                ("mixed_f32", dict(precision=la.PREC_MIXED_F32, inner_tol=1e-5, **S)),
                ("mixed_bf16_matrix_f32_factors",
                 dict(precision=la.PREC_MIXED_BF16_F32P, inner_tol=1e-3, **S)),
                ("fp64_jacobi6", dict(precision=la.PREC_FP64, **J)),
                ("mixed_f32_jacobi6", dict(precision=la.PREC_MIXED_F32, inner_tol=1e-4, **J)),
                ("mixed_bf16_matrix_f32_factors_jacobi6",
                 dict(precision=la.PREC_MIXED_BF16_F32P, inner_tol=1e-3, **J))]
    xs = {}
    for name, kw in variants:
        x = torch.zeros(A.n, dtype=torch.float64, device=dev)
        res = la.fgmres(dA, P, b, x, rel_tol=1e-10, restart=30, max_iters=3000, **kw)
        xh = x.cpu().numpy()
        rr = bh - orc.spmv(Ao, xh, alg="parallel", nthreads=orc.max_threads())
        true_rel = np.linalg.norm(rr) / np.linalg.norm(bh)
        ref = xs.get("fp64_jacobi6" if "jacobi" in name else "fp64", xh)
        dx = np.linalg.norm(xh - ref) / np.linalg.norm(ref)
        print("c4 gmres", name, res, "host true residual %.3e, x vs fp64 %.3e" % (true_rel, dx),
              flush=True)
        assert res["converged"] and res["rel_residual"] < 1e-10
        assert true_rel <= 1.01e-10
        assert dx <= 1e-6
        xs[name] = xh
