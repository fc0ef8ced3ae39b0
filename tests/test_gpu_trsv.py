"""GPU triangular solves and ILU(k)/IC(k) apply against the oracle's
ForwardSubstitution / BackwardSubstitution / ForwardSubstitutionT
(sparse_mat_op/triangle_solve.hpp:32-98) and the preconditioner applies
(incomplete_lu.cpp:36-48, incomplete_cholesky.cpp:21-29).

The reference checks its level-scheduled and optimized solves against the
serial ones at 1e-14 element-wise and repeats them 100 times to expose races
(tests/triangular_solve_test.cpp:207-347). Here: FP64 element-wise
|x - ref| <= 1e-12 |ref| + 1e-14 max|ref| (FMA and division may round
differently), every mode repeated 5 times with bitwise-equal results.
FP32 / BF16 factors are checked against the FP64 oracle at 1e-4 / 5e-2
relative (norm-wise) on well-conditioned factors.
"""
import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import as_lib_csr, bf16_rne, orc, random_box_triangle, read_mtx

pytestmark = pytest.mark.gpu

MODES = [la.TRSV_LEVEL, la.TRSV_SYNCFREE]


def _o(A):
    return orc.CSR(A.n, A.ai, A.aj, A.av, ncols=A.ncols)


def _factor_cases():
    """(name, combined ILU factor) -- the reference's tests factor with
    incomplete_lu_k level 5 (tests/triangular_solve_test.cpp:84-87)."""
    out = []
    A = read_mtx("ex5.mtx")
    out.append(("ex5_ilu5", lambda: orc.iluk(A, 5)[0]))
    out.append(("ex5_ilu5_base1", lambda: orc.iluk(read_mtx("ex5.mtx", base=1), 5)[0]))
    out.append(("poisson2d_48_ilu0", lambda: orc.iluk(_o(la.gen.poisson2d(48)), 0)[0]))
    out.append(("laplace3d_14_ilu1", lambda: orc.iluk(_o(la.gen.laplace3d(14)), 1)[0]))
    out.append(("aniso27_8_ilu1", lambda: orc.iluk(_o(la.gen.aniso27(8)), 1)[0]))
    out.append(("powerlaw_ilu0", lambda: orc.iluk(_o(la.gen.powerlaw(3000, 8.0, 1.0, 400, seed=2)), 0)[0]))
    return out


def _close(got, ref, tol=1e-12):
    m = np.abs(ref).max() if ref.size else 0.0
    return np.all(np.abs(got - ref) <= tol * np.abs(ref) + 1e-14 * m)


@pytest.mark.parametrize("name,make", _factor_cases(), ids=[c[0] for c in _factor_cases()])
@pytest.mark.parametrize("mode", MODES)
def test_forward_backward_f64(dev, name, make, mode):
    import torch
    F = make()
    L, D, U = orc.split_ldu(F)
    print("case", name, "mode", mode, "n", F.n, flush=True)
    b = np.ones(F.n) if "ex5" in name else la.gen.uniform(F.n, 3)
    bt = torch.tensor(b, device=dev)
    fw = la.TriangularSolve(as_lib_csr(L), None, upper=False)
    bw = la.TriangularSolve(as_lib_csr(U), D, upper=True)
    ref_f = orc.forward_sub(L, b)
    ref_b = orc.backward_sub(U, D, b)
    first_f = first_b = None
    for rep in range(5):
        x = torch.zeros(F.n, dtype=torch.float64, device=dev)
        fw.solve(bt, x, mode=mode)
        got = x.cpu().numpy()
        assert _close(got, ref_f), (name, "forward", rep)
        if first_f is None:
            first_f = got
        np.testing.assert_array_equal(got, first_f)
        x = torch.zeros(F.n, dtype=torch.float64, device=dev)
        bw.solve(bt, x, mode=mode)
        got = x.cpu().numpy()
        assert _close(got, ref_b), (name, "backward", rep)
        if first_b is None:
            first_b = got
        np.testing.assert_array_equal(got, first_b)


@pytest.mark.parametrize("mode", MODES)
def test_level_counts_match_topological_sort(dev, mode):
    """The plan's level count equals TopologicalSort2's (matrix_utils.hpp:637);
    for a natural-order 3-D 7-point grid L factor it is nx+ny+nz-2."""
    F = orc.iluk(_o(la.gen.laplace3d(9, 7, 5)), 0)[0]
    L, D, U = orc.split_ldu(F)
    T = la.TriangularSolve(as_lib_csr(L), None, upper=False)
    assert T.info()["levels"] == 9 + 7 + 5 - 2 == orc.topo_sort2(L, False)[2]


@pytest.mark.parametrize("name,make", _factor_cases()[2:], ids=[c[0] for c in _factor_cases()[2:]])
@pytest.mark.parametrize("mode", MODES)
def test_transposed_solve(dev, name, make, mode):
    """(U + D)^T x = b and (I + L)^T x = b against dense solves; for a unit
    strict U the oracle's ForwardSubstitutionT (triangle_solve.hpp:87-98)."""
    import torch
    F = make()
    L, D, U = orc.split_ldu(F)
    b = la.gen.uniform(F.n, 5)
    bt = torch.tensor(b, device=dev)
    bw = la.TriangularSolve(as_lib_csr(U), D, upper=True)
    x = torch.zeros(F.n, dtype=torch.float64, device=dev)
    bw.solve_transpose(bt, x, mode=mode)
    Ud = U.to_scipy().toarray() + np.diag(D)
    ref = np.linalg.solve(Ud.T, b)
    assert np.linalg.norm(x.cpu().numpy() - ref) <= 1e-10 * np.linalg.norm(ref)
    uw = la.TriangularSolve(as_lib_csr(U), None, upper=True)
    uw.solve_transpose(bt, x, mode=mode)
    assert _close(x.cpu().numpy(), orc.forward_sub_t(U, b), 1e-11)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name,make", _factor_cases()[2:], ids=[c[0] for c in _factor_cases()[2:]])
def test_ilu_apply(dev, name, make, mode):
    import torch
    F = make()
    A = F  # the apply only needs the factor
    b = la.gen.uniform(F.n, 6)
    P = la.Preconditioner(as_lib_csr(F), "ilu")
    z = torch.zeros(F.n, dtype=torch.float64, device=dev)
    P.apply(torch.tensor(b, device=dev), z, mode=mode)
    ref = orc.prec_apply(A, "ilu", F, b)
    assert _close(z.cpu().numpy(), ref, 1e-11)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("lvl", [0, 1, 2])
def test_ic_apply(dev, mode, lvl):
    import torch
    A = _o(la.gen.laplace3d(12))
    U = orc.split_triangle(A, True)
    R, ok = orc.icc(U, lvl)
    assert ok
    b = la.gen.uniform(A.n, 6)
    P = la.Preconditioner(as_lib_csr(R), "ic")
    z = torch.zeros(A.n, dtype=torch.float64, device=dev)
    P.apply(torch.tensor(b, device=dev), z, mode=mode)
    ref = orc.prec_apply(A, "ic", R, b)
    assert _close(z.cpu().numpy(), ref, 1e-11)
    zt = torch.zeros(A.n, dtype=torch.float64, device=dev)
    P.apply_transpose(torch.tensor(b, device=dev), zt, mode=mode)
    assert _close(zt.cpu().numpy(), ref, 1e-11)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("lvl", [0, 2])
def test_ic_k_additive_apply(dev, mode, lvl):
    """The incomplete_cholesky_k factor (additive shift, its own row order;
    incomplete_cholesky.cpp:195-274) made by the product's host phase, then
    applied on the device as U^-1 U^-T (:21-29)."""
    import torch
    A = _o(la.gen.laplace3d(12))
    U = orc.split_triangle(A, True)
    R, s = la.ick_shifted(as_lib_csr(U), lvl)
    Ro, ok = orc.ick(U, lvl, shift=s)
    assert ok and s == 0.0
    np.testing.assert_array_equal(R.av, Ro.av)
    b = la.gen.uniform(A.n, 8)
    P = la.Preconditioner(R, "ic")
    z = torch.zeros(A.n, dtype=torch.float64, device=dev)
    P.apply(torch.tensor(b, device=dev), z, mode=mode)
    ref = orc.prec_apply(A, "ic", Ro, b)
    assert _close(z.cpu().numpy(), ref, 1e-11)


@pytest.mark.parametrize("mode", MODES)
def test_ilu_apply_transpose(dev, mode):
    import torch
    F = orc.iluk(_o(la.gen.powerlaw(2000, 8.0, 1.0, 300, seed=4)), 1)[0]
    L, D, U = orc.split_ldu(F)
    Ld = L.to_scipy().toarray() + np.eye(F.n)
    Ud = U.to_scipy().toarray() + np.diag(D)
    b = la.gen.uniform(F.n, 8)
    P = la.Preconditioner(as_lib_csr(F), "ilu")
    z = torch.zeros(F.n, dtype=torch.float64, device=dev)
    P.apply_transpose(torch.tensor(b, device=dev), z, mode=mode)
    ref = np.linalg.solve((Ld @ Ud).T, b)
    assert np.linalg.norm(z.cpu().numpy() - ref) <= 1e-10 * np.linalg.norm(ref)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("val,tol", [(la.F32, 1e-4), (la.BF16, 5e-2)])
def test_lowprec_apply(dev, mode, val, tol):
    """IC(0) apply with FP32 / BF16 factors, FP32 arithmetic. Against the
    oracle's FP64 apply with the factors rounded by the tests themselves
    (off-diagonal values, and the reciprocal diagonal the plans store):
    only the FP32 arithmetic differs, 1e-5 relative. Against the unrounded
    FP64 apply: the storage error, carried through both triangles."""
    import torch
    A = _o(la.gen.laplace3d(16))
    U = orc.split_triangle(A, True)
    R, ok = orc.icc(U, 0)
    b = la.gen.uniform(A.n, 6)
    ref = orc.prec_apply(A, "ic", R, b)
    rnd = (lambda v: v.astype(np.float32).astype(np.float64)) if val == la.F32 else bf16_rne
    rows = np.repeat(np.arange(R.n), np.diff(R.ai))
    on_diag = rows == R.aj
    av = rnd(R.av)
    av[on_diag] = 1.0 / rnd(1.0 / R.av[on_diag])
    ref_r = orc.prec_apply(A, "ic", orc.CSR(R.n, R.ai, R.aj, av), b)
    P = la.Preconditioner(as_lib_csr(R), "ic")
    for vdt in (torch.float32, torch.float64):
        z = torch.zeros(A.n, dtype=vdt, device=dev)
        P.apply(torch.tensor(b, device=dev, dtype=vdt), z, mode=mode, val=val)
        zh = z.double().cpu().numpy()
        err = np.linalg.norm(zh - ref) / np.linalg.norm(ref)
        err_r = np.linalg.norm(zh - ref_r) / np.linalg.norm(ref_r)
        print("val", val, "vec", vdt, "err vs fp64 %.2e, vs rounded factors %.2e" % (err, err_r),
              flush=True)
        assert err_r < 1e-5, (val, vdt, err_r)
        assert err < tol, (val, vdt, err)


def test_rejects_non_strict_triangle():
    """TopologicalSort2 needs strict triangles (matrix_utils.hpp:654-656);
    the plan refuses a diagonal or wrong-side entry instead of mis-solving."""
    A = _o(la.gen.poisson2d(6))
    Lw = orc.split_triangle(A, False)  # keeps the diagonal
    with pytest.raises(la.LaError):
        la.TriangularSolve(as_lib_csr(Lw), None, upper=False)


# ------------------------------------------------------------- tiled schedule

def _grid_cases():
    """Triangles whose pattern is a grid (the tiled sync-free schedule's
    case): natural and RCM order, 2-D and 3-D, IC(0) and ILU(0)."""
    def rcm(A):
        ip, p = orc.rcm(A)
        return orc.permute(A, ip, p)
    return [("poisson2d_96_natural", lambda: _o(la.gen.poisson2d(96))),
            ("poisson2d_80_rcm", lambda: rcm(_o(la.gen.poisson2d(80)))),
            ("laplace3d_24_rcm", lambda: rcm(_o(la.gen.laplace3d(24)))),
            ("laplace3d_20x16x12_natural", lambda: _o(la.gen.laplace3d(20, 16, 12)))]


@pytest.mark.parametrize("name,make", _grid_cases(), ids=[c[0] for c in _grid_cases()])
def test_tiled_schedule_ic0_apply(dev, name, make):
    """The sync-free mode runs the tiled schedule on grid patterns (both
    triangles, any row order) and gives the oracle's apply to 1e-12 in FP64,
    the level-launch result to 1e-13, and the FP32 apply to the FP32
    bound; also transposed (A symmetric)."""
    import torch
    A = make()
    U = orc.split_triangle(A, True)
    R, ok = orc.icc(U, 0)
    P = la.Preconditioner(as_lib_csr(R), "ic")
    tl, tu = P.tiles()
    print("case", name, "n", A.n, "tiles", tl, tu, flush=True)
    assert tl > 1 and tu > 1
    b = la.gen.uniform(A.n, 12)
    ref = orc.prec_apply(A, "ic", R, b)
    bt = torch.tensor(b, device=dev)
    z = torch.zeros_like(bt)
    P.apply(bt, z, mode=la.TRSV_SYNCFREE)
    assert _close(z.cpu().numpy(), ref, 1e-12)
    zl = torch.zeros_like(bt)
    P.apply(bt, zl, mode=la.TRSV_LEVEL)
    assert _close(z.cpu().numpy(), zl.cpu().numpy(), 1e-13)
    for _ in range(3):  # repeated solves reuse the plan's control words
        z2 = torch.zeros_like(bt)
        P.apply(bt, z2, mode=la.TRSV_SYNCFREE)
        assert torch.equal(z2, z)
    z32 = torch.zeros(A.n, dtype=torch.float32, device=dev)
    P.apply(bt.float(), z32, mode=la.TRSV_SYNCFREE, val=la.F32)
    assert np.linalg.norm(z32.double().cpu().numpy() - ref) <= 1e-4 * np.linalg.norm(ref)


@pytest.mark.parametrize("name,make", _grid_cases()[:3], ids=[c[0] for c in _grid_cases()[:3]])
def test_tiled_schedule_ilu0_unit_lower(dev, name, make):
    """ILU(0) on a grid: the unit lower triangle and the upper one with its
    diagonal both tile; forward and backward solves match the oracle."""
    import torch
    F = orc.iluk(make(), 0)[0]
    L, D, U = orc.split_ldu(F)
    fw = la.TriangularSolve(as_lib_csr(L), None, upper=False)
    bw = la.TriangularSolve(as_lib_csr(U), D, upper=True)
    assert fw.info()["tiled"] and bw.info()["tiled"]
    b = la.gen.uniform(F.n, 13)
    bt = torch.tensor(b, device=dev)
    x = torch.zeros_like(bt)
    fw.solve(bt, x, mode=la.TRSV_SYNCFREE)
    assert _close(x.cpu().numpy(), orc.forward_sub(L, b), 1e-12)
    bw.solve(bt, x, mode=la.TRSV_SYNCFREE)
    assert _close(x.cpu().numpy(), orc.backward_sub(U, D, b), 1e-12)


def _star13(nx, ny, nz):
    """3-D 13-point star stencil (neighbours at distance 1 and 2 along each
    axis), diagonally dominant: its IC(0) triangle has rows of up to 6
    entries, the tiled kernel's 8-entry form."""
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [np.full(n, 10.0)]
    for ax in range(3):
        for d, w in ((1, -1.0), (2, -0.5)):
            sa, sb = [slice(None)] * 3, [slice(None)] * 3
            sa[ax], sb[ax] = slice(0, idx.shape[ax] - d), slice(d, None)
            a, b = idx[tuple(sa)].ravel(), idx[tuple(sb)].ravel()
            rows += [a, b]
            cols += [b, a]
            vals += [np.full(a.size, w)] * 2
    r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    ai = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=n))]).astype(np.int32)
    return orc.CSR(n, ai, c.astype(np.int32), v)


@pytest.mark.parametrize("dims", [(24, 20, 16), (17, 13, 11)])
def test_tiled_schedule_wide_rows(dev, dims):
    """Rows of 5-8 entries run the tiled kernel's 8-entry form (a shallower
    register pipeline); both triangles of the IC(0) apply match the oracle
    in FP64 and the FP32 bound, and repeat bitwise."""
    import torch
    A = _star13(*dims)
    R, ok = orc.icc(orc.split_triangle(A, True), 0)
    assert ok
    P = la.Preconditioner(as_lib_csr(R), "ic")
    tl, tu = P.tiles()
    print("case star13", dims, "tiles", tl, tu, flush=True)
    assert tl > 1 and tu > 1
    b = la.gen.uniform(A.n, 15)
    ref = orc.prec_apply(A, "ic", R, b)
    bt = torch.tensor(b, device=dev)
    z = torch.zeros_like(bt)
    P.apply(bt, z, mode=la.TRSV_SYNCFREE)
    assert _close(z.cpu().numpy(), ref, 1e-12)
    z2 = torch.zeros_like(bt)
    P.apply(bt, z2, mode=la.TRSV_SYNCFREE)
    assert torch.equal(z2, z)
    z32 = torch.zeros(A.n, dtype=torch.float32, device=dev)
    P.apply(bt.float(), z32, mode=la.TRSV_SYNCFREE, val=la.F32)
    assert np.linalg.norm(z32.double().cpu().numpy() - ref) <= 1e-4 * np.linalg.norm(ref)


def _box_cases():
    """Natural-order grids whose ILU factors have rows of 9-32 entries:
    27-point ILU(1) (several tiles per plane at 40 x 70 x 6), 27-point
    ILU(0) and 7-point ILU(2)."""
    return [("aniso27_24x20x16_ilu1", lambda: orc.iluk(_o(la.gen.aniso27(24, 20, 16)), 1)[0]),
            ("aniso27_17x13x11_ilu1", lambda: orc.iluk(_o(la.gen.aniso27(17, 13, 11)), 1)[0]),
            ("aniso27_40x70x6_ilu1", lambda: orc.iluk(_o(la.gen.aniso27(40, 70, 6)), 1)[0]),
            ("aniso27_18x14x10_ilu0", lambda: orc.iluk(_o(la.gen.aniso27(18, 14, 10)), 0)[0]),
            ("laplace3d_20x16x12_ilu2", lambda: orc.iluk(_o(la.gen.laplace3d(20, 16, 12)), 2)[0])]


@pytest.mark.parametrize("name,make", _box_cases(), ids=[c[0] for c in _box_cases()])
def test_tiled_schedule_box_grid(dev, name, make):
    """ILU factors of natural-order grids with rows of up to 32 entries run
    the tiled kernel's four-lane form on tiles of 8 x 2 lines skewed in
    (j + 2k, k), the box grid recognised from the row indices. Both
    triangles match the oracle in FP64 (1e-12), the level launches (1e-13)
    and repeat bitwise; the FP32 apply stays within 1e-5 of the FP64 one
    (its storage rounding gives about 2e-7 here)."""
    import torch
    F = make()
    P = la.Preconditioner(as_lib_csr(F), "ilu")
    tl, tu = P.tiles()
    print("case", name, "tiles", tl, tu, flush=True)
    assert tl > 1 and tu > 1
    b = la.gen.uniform(F.n, 16)
    ref = orc.prec_apply(F, "ilu", F, b)
    bt = torch.tensor(b, device=dev)
    z = torch.zeros_like(bt)
    P.apply(bt, z, mode=la.TRSV_SYNCFREE)
    assert _close(z.cpu().numpy(), ref, 1e-12)
    zl = torch.zeros_like(bt)
    P.apply(bt, zl, mode=la.TRSV_LEVEL)
    assert _close(z.cpu().numpy(), zl.cpu().numpy(), 1e-13)
    for _ in range(3):
        z2 = torch.zeros_like(bt)
        P.apply(bt, z2, mode=la.TRSV_SYNCFREE)
        assert torch.equal(z2, z)
    P.add_dtype(la.F32)
    z32 = torch.zeros(F.n, dtype=torch.float32, device=dev)
    P.apply(bt.float(), z32, mode=la.TRSV_SYNCFREE, val=la.F32)
    assert np.linalg.norm(z32.double().cpu().numpy() - ref) <= 1e-5 * np.linalg.norm(ref)
    L, D, U = orc.split_ldu(F)
    fw = la.TriangularSolve(as_lib_csr(L), None, upper=False)
    bw = la.TriangularSolve(as_lib_csr(U), D, upper=True)
    assert fw.info()["tiled"] and bw.info()["tiled"]
    x = torch.zeros_like(bt)
    fw.solve(bt, x, mode=la.TRSV_SYNCFREE)
    assert _close(x.cpu().numpy(), orc.forward_sub(L, b), 1e-12)
    bw.solve(bt, x, mode=la.TRSV_SYNCFREE)
    assert _close(x.cpu().numpy(), orc.backward_sub(U, D, b), 1e-12)


def test_tiled_schedule_box_grid_base1(dev):
    """The box grid is recognised the same way from 1-based indices: both
    triangles of a 1-based copy run tiled and give the 0-based solves'
    results bit for bit."""
    import torch
    F = orc.iluk(_o(la.gen.aniso27(17, 13, 11)), 1)[0]
    L, D, U = orc.split_ldu(F)
    b = la.gen.uniform(F.n, 18)
    bt = torch.tensor(b, device=dev)
    for T, diag, upper in ((L, None, False), (U, D, True)):
        A0 = as_lib_csr(T)
        A1 = la.CSR(T.n, np.asarray(T.ai) + 1, np.asarray(T.aj) + 1, T.av, ncols=T.ncols)
        s0 = la.TriangularSolve(A0, diag, upper=upper)
        s1 = la.TriangularSolve(A1, diag, upper=upper)
        print("case aniso27_17x13x11_ilu1 base1", upper, s1.info(), flush=True)
        assert s0.info()["tiled"] and s1.info()["tiled"]
        assert s1.info()["tiles"] == s0.info()["tiles"]
        x0, x1 = torch.zeros_like(bt), torch.zeros_like(bt)
        s0.solve(bt, x0, mode=la.TRSV_SYNCFREE)
        s1.solve(bt, x1, mode=la.TRSV_SYNCFREE)
        assert torch.equal(x0, x1)
        ref = orc.backward_sub(U, D, b) if upper else orc.forward_sub(L, b)
        assert _close(x1.cpu().numpy(), ref, 1e-12)


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("upper", [False, True])
def test_tiled_schedule_random_box_patterns(dev, seed, upper):
    """Box-grid triangles whose rows hold random subsets of the 27-point
    ILU(1) pattern (9-31 entries, random values, random diagonal) run the
    four-lane tiled kernel and match a direct triangular solve to 1e-12,
    the level launches to 1e-13, and repeat bitwise."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    import torch
    T = random_box_triangle(13, 11, 9, seed, upper)
    D = 1.0 + np.random.default_rng(seed).random(T.n)
    s = la.TriangularSolve(as_lib_csr(T), D, upper=upper)
    print("case random_box", seed, upper, s.info(), flush=True)
    assert s.info()["tiled"] and s.info()["tiles"] > 1
    b = la.gen.uniform(T.n, 24)
    M = sp.csr_matrix((T.av, T.aj, T.ai), shape=(T.n, T.n)) + sp.diags(D)
    ref = spl.spsolve_triangular(M.tocsr(), b, lower=not upper)
    bt = torch.tensor(b, device=dev)
    x = torch.zeros_like(bt)
    s.solve(bt, x, mode=la.TRSV_SYNCFREE)
    assert _close(x.cpu().numpy(), ref, 1e-12)
    xl = torch.zeros_like(bt)
    s.solve(bt, xl, mode=la.TRSV_LEVEL)
    assert _close(x.cpu().numpy(), xl.cpu().numpy(), 1e-13)
    for _ in range(3):
        x2 = torch.zeros_like(bt)
        s.solve(bt, x2, mode=la.TRSV_SYNCFREE)
        assert torch.equal(x2, x)


def test_box_grid_rows_too_long_keep_level_chunks(dev):
    """A natural-order box grid whose factor has rows longer than the tiled
    kernel's four lanes take (27-point ILU(2): up to 57 entries) is refused
    by the tiled layout; both triangles keep the level-chunk schedule and
    solve exactly (oracle at 1e-12, the level launches at 1e-13, bitwise
    repeats)."""
    import torch
    F = orc.iluk(_o(la.gen.aniso27(16, 14, 12)), 2)[0]
    L, D, U = orc.split_ldu(F)
    assert max(np.diff(np.asarray(L.ai)).max(), np.diff(np.asarray(U.ai)).max()) > 32
    fw = la.TriangularSolve(as_lib_csr(L), None, upper=False)
    bw = la.TriangularSolve(as_lib_csr(U), D, upper=True)
    print("case aniso27_16x14x12_ilu2", fw.info(), bw.info(), flush=True)
    assert not fw.info()["tiled"] and not bw.info()["tiled"]
    b = la.gen.uniform(F.n, 17)
    bt = torch.tensor(b, device=dev)
    for T, ref in ((fw, orc.forward_sub(L, b)), (bw, orc.backward_sub(U, D, b))):
        x = torch.zeros_like(bt)
        T.solve(bt, x, mode=la.TRSV_SYNCFREE)
        assert _close(x.cpu().numpy(), ref, 1e-12)
        xl = torch.zeros_like(bt)
        T.solve(bt, xl, mode=la.TRSV_LEVEL)
        assert _close(x.cpu().numpy(), xl.cpu().numpy(), 1e-13)
        for _ in range(3):
            x2 = torch.zeros_like(bt)
            T.solve(bt, x2, mode=la.TRSV_SYNCFREE)
            assert torch.equal(x2, x)


def test_tiled_schedule_off_and_non_grid(dev, monkeypatch):
    """LAAMD_SF_TILED=0 at plan creation keeps the level-chunk schedule;
    a pattern that is not a grid keeps it too; both solve exactly."""
    import torch
    A = _o(la.gen.laplace3d(16))
    R, ok = orc.icc(orc.split_triangle(A, True), 0)
    monkeypatch.setenv("LAAMD_SF_TILED", "0")
    P0 = la.Preconditioner(as_lib_csr(R), "ic")
    monkeypatch.delenv("LAAMD_SF_TILED")
    P1 = la.Preconditioner(as_lib_csr(R), "ic")
    assert P0.tiles() == (0, 0) and min(P1.tiles()) > 0
    b = torch.tensor(la.gen.uniform(A.n, 14), device=dev)
    z0, z1 = torch.zeros_like(b), torch.zeros_like(b)
    P0.apply(b, z0, mode=la.TRSV_SYNCFREE)
    P1.apply(b, z1, mode=la.TRSV_SYNCFREE)
    assert _close(z1.cpu().numpy(), z0.cpu().numpy(), 1e-13)
    F = orc.iluk(_o(la.gen.powerlaw(3000, 8.0, 1.0, 400, seed=2)), 0)[0]
    L, D, U = orc.split_ldu(F)
    fw = la.TriangularSolve(as_lib_csr(L), None, upper=False)
    assert not fw.info()["tiled"]
