"""GPU SpMV parity (through the C ABI) against the oracle's SerialSPMV
restatement (sparse_mat_op/spmv.hpp:12-29).

Tolerances: FP64 values/vectors/accumulation -- norm-wise relative error
<= 1e-12 and element-wise |y - ref| <= 1e-12 * (|alpha| |A| |x| + |beta| |y0|)_i
(only summation order / FMA rounding may differ). FP32 / BF16 storage is
checked twice: against the oracle run on the values rounded to the storage
type by the tests' own round-to-nearest-even (only the FP32 accumulation
and the FP32 vector may differ: <= 2e-6 of (|A| |x|)_i), and against the
FP64 oracle with the storage rounding bound (unit roundoff 2^-24 / 2^-9,
doubled, plus the accumulation term): <= 2e-6 (FP32) and <= 2^-8 + 2e-6
(BF16) of (|A| |x|)_i. The device FP64 -> BF16 conversion is checked bit
for bit on a diagonal matrix.
"""
import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import (as_lib_csr, bf16_bits, bf16_rne, bf16_special_values, from_dense, orc,
                     random_sparse, read_mtx)

pytestmark = pytest.mark.gpu


def _cases():
    cs = [("ex5", lambda: read_mtx("ex5.mtx")),
          ("rdist1", lambda: read_mtx("rdist1.mtx")),
          ("rdist1_base1", lambda: read_mtx("rdist1.mtx", base=1)),
          ("poisson2d_64", lambda: _o(la.gen.poisson2d(64))),
          ("laplace3d_16", lambda: _o(la.gen.laplace3d(16))),
          ("powerlaw_longrows", lambda: _o(la.gen.powerlaw(20000, 16.0, 2.0, 6000, seed=5))),
          ("aniso27_10", lambda: _o(la.gen.aniso27(10, 9, 8))),
          ("empty_rows", lambda: from_dense(_empty_rows())),
          ("one_long_row", lambda: from_dense(_one_long_row())),
          ("all_empty", lambda: orc.CSR(7, np.zeros(8), [], [])),
          ("one_by_one", lambda: from_dense([[3.0]])),
          ("K_bin_base1", lambda: _bin("K.bin")),
          ("G2_bin_rect", lambda: _bin("G2.bin")),
          # narrow rows (at most 4 per row: two slices per wavefront), with
          # an odd slice count and a ragged last slice
          ("lower_laplace3d_12", lambda: orc.split_triangle(_o(la.gen.laplace3d(12)), False)),
          ("bidiag_273", lambda: _bidiag(273)),
          # sliced ELL with 16-bit column offsets at their limits (+-32767),
          # and one whose offsets do not fit (int32 columns)
          ("offsets_at_16bit_limit", lambda: _band(70000, 32767)),
          ("offsets_beyond_16bit", lambda: _band(70000, 40000))]
    return cs


def _band(n, off):
    """rows with columns i - off, i, i + off where they exist"""
    rows, cols = [], []
    for d in (-off, 0, off):
        i = np.arange(max(0, -d), min(n, n - d))
        rows.append(i)
        cols.append(i + d)
    r, c = np.concatenate(rows), np.concatenate(cols)
    o = np.lexsort((c, r))
    r, c = r[o], c[o]
    v = np.random.default_rng(off).standard_normal(len(r))
    ai = np.zeros(n + 1, np.int64)
    np.add.at(ai, r + 1, 1)
    return orc.CSR(n, np.cumsum(ai).astype(np.int32), c.astype(np.int32), v)


def _bidiag(n):
    D = np.diag(np.full(n, 2.0)) + np.diag(np.full(n - 1, -1.0), -1)
    D[5, :] = 0.0  # an empty row
    return from_dense(D)


def _o(A):
    return orc.CSR(A.n, A.ai, A.aj, A.av, ncols=A.ncols)


def _bin(name):
    """the reference's binary COO fixtures, read 1-based as its tests do"""
    import os
    from helpers import GOLDEN
    return _o(la.read_binary_coo(os.path.join(GOLDEN, name), base=1))


def _empty_rows():
    D = random_sparse(300, 0.05, seed=1)
    D[::7] = 0.0
    D[-5:] = 0.0
    D[:3] = 0.0
    return D


def _one_long_row():
    D = random_sparse(50, 0.02, seed=2, ncols=3000)
    D[17] = np.random.default_rng(3).standard_normal(3000)
    return D


def _absbound(A, x):
    B = orc.CSR(A.n, A.ai, A.aj, np.abs(A.av), ncols=A.ncols)
    return orc.spmv(B, np.abs(x))


@pytest.mark.parametrize("name,make", _cases(), ids=[c[0] for c in _cases()])
@pytest.mark.parametrize("alg", [la.SPMV_ROW, la.SPMV_SEGSUM, la.SPMV_ALBUS, la.SPMV_AUTO,
                                 la.SPMV_SELL])
def test_spmv_f64(dev, name, make, alg):
    import torch
    A = make()
    print("case", name, "alg", alg, "n", A.n, "nnz", A.nnz, flush=True)
    x = la.gen.uniform(A.ncols, 7)
    y0 = la.gen.uniform(A.n, 8)
    dA = la.DeviceMatrix(as_lib_csr(A))
    xt = torch.tensor(x, device=dev)
    for alpha, beta in [(1.0, 0.0), (2.5, -0.5), (-1.0, 1.0)]:
        yt = torch.tensor(y0, device=dev)
        dA.spmv(xt, yt, alpha=alpha, beta=beta, alg=alg)
        ref = orc.spmv(A, x, y0, alpha=alpha, beta=beta)
        got = yt.cpu().numpy()
        scale = abs(alpha) * _absbound(A, x) + abs(beta) * np.abs(y0)
        assert np.all(np.abs(got - ref) <= 1e-12 * scale + 1e-300), (name, alpha, beta)
        if np.linalg.norm(ref) > 0:
            assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-12 or \
                np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(scale)


@pytest.mark.parametrize("alg", [la.SPMV_ROW, la.SPMV_ALBUS])
def test_spmv_beta_zero_ignores_y(dev, alg):
    """beta == 0 never reads y (SerialSPMV, spmv.hpp:26): NaN in y stays out."""
    import torch
    A = _o(la.gen.poisson2d(40))
    x = la.gen.uniform(A.n, 9)
    dA = la.DeviceMatrix(as_lib_csr(A))
    y = torch.full((A.n,), float("nan"), dtype=torch.float64, device=dev)
    dA.spmv(torch.tensor(x, device=dev), y, alg=alg)
    assert torch.isfinite(y).all()
    np.testing.assert_allclose(y.cpu().numpy(), orc.spmv(A, x), rtol=0, atol=1e-13)


BF16_BOUND = 2.0 ** -8 + 2e-6   # twice BF16's unit roundoff + FP32 accumulation
ACC32_BOUND = 2e-6              # FP32 accumulation and FP32 vectors


def _round_values(A, val):
    av = A.av.astype(np.float32).astype(np.float64) if val == la.F32 else bf16_rne(A.av)
    return orc.CSR(A.n, A.ai, A.aj, av, ncols=A.ncols)


@pytest.mark.parametrize("name,make", _cases()[:7], ids=[c[0] for c in _cases()[:7]])
@pytest.mark.parametrize("val,bound", [(la.F32, 2e-6), (la.BF16, BF16_BOUND)])
@pytest.mark.parametrize("alg", [la.SPMV_ROW, la.SPMV_ALBUS, la.SPMV_SELL])
def test_spmv_lowprec(dev, name, make, val, bound, alg):
    import torch
    A = make()
    print("case", name, "val", val, "alg", alg, flush=True)
    x = la.gen.uniform(A.ncols, 7)
    dA = la.DeviceMatrix(as_lib_csr(A))
    ref = orc.spmv(A, x)
    ref_rounded = orc.spmv(_round_values(A, val), x)
    scale = _absbound(A, x)
    for vec in (torch.float32, torch.float64):
        xt = torch.tensor(x, device=dev, dtype=vec)
        yt = torch.zeros(A.n, device=dev, dtype=vec)
        dA.spmv(xt, yt, alg=alg, val=val, acc=la.F32)
        got = yt.double().cpu().numpy()
        assert np.all(np.abs(got - ref_rounded) <= ACC32_BOUND * scale + 1e-30), (name, vec)
        assert np.all(np.abs(got - ref) <= bound * scale + 1e-30), (name, vec)


@pytest.mark.parametrize("alg", [la.SPMV_ROW, la.SPMV_ALBUS, la.SPMV_SELL])
def test_bf16_conversion_bitwise(dev, alg):
    """The device FP64 -> BF16 conversion (la_mat_add_dtype) rounds to
    nearest even in one step: a diagonal matrix times x = 1 in FP32 reads
    back bf16(a_ii) exactly; ties, near-ties, subnormals, overflow, inf and
    NaN are compared with the tests' integer-valued rounding."""
    import torch
    v = bf16_special_values()
    v = np.concatenate([v, [np.nan, -np.nan]])
    n = v.size
    A = orc.CSR(n, np.arange(n + 1), np.arange(n), v)
    dA = la.DeviceMatrix(as_lib_csr(A))
    dA.add_dtype(la.BF16)
    y = torch.zeros(n, device=dev, dtype=torch.float32)
    dA.spmv(torch.ones(n, device=dev, dtype=torch.float32), y, alg=alg, val=la.BF16,
            acc=la.F32)
    got = y.cpu().numpy()
    want = bf16_rne(v)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    fin = np.isfinite(want)
    gb = got.view(np.uint32)
    assert np.all(gb[fin] & 0xFFFF == 0)
    nz = fin & (want != 0)
    assert np.array_equal((gb[nz] >> 16).astype(np.uint16), bf16_bits(want[nz])), \
        v[nz][(gb[nz] >> 16).astype(np.uint16) != bf16_bits(want[nz])]
    assert np.all(got[want == 0] == 0)
    inf = np.isinf(want)
    assert np.array_equal(got[inf], want[inf].astype(np.float32))


def test_spmv_f64_storage_f32_accumulate(dev):
    import torch
    A = _o(la.gen.powerlaw(5000, 16.0, 1.0, 20000, seed=9))
    x = la.gen.uniform(A.n, 7)
    dA = la.DeviceMatrix(as_lib_csr(A))
    yt = torch.zeros(A.n, device=dev, dtype=torch.float64)
    dA.spmv(torch.tensor(x, device=dev), yt, val=la.F64, acc=la.F32)
    ref = orc.spmv(A, x)
    assert np.all(np.abs(yt.cpu().numpy() - ref) <= 2e-6 * _absbound(A, x))


@pytest.mark.parametrize("name,make", _cases()[:3] + _cases()[6:8],
                         ids=[c[0] for c in _cases()[:3] + _cases()[6:8]])
def test_spmv_transpose(dev, name, make):
    import torch
    A = make()
    At = orc.transpose(A)
    x = la.gen.uniform(A.n, 3)
    y0 = la.gen.uniform(A.ncols, 4)
    dA = la.DeviceMatrix(as_lib_csr(A))
    yt = torch.tensor(y0, device=dev)
    dA.spmv_transpose(torch.tensor(x, device=dev), yt, alpha=1.5, beta=0.25)
    ref = orc.spmv(At, x, y0, alpha=1.5, beta=0.25)
    scale = 1.5 * _absbound(At, x) + 0.25 * np.abs(y0)
    assert np.all(np.abs(yt.cpu().numpy() - ref) <= 1e-12 * scale + 1e-300)


def test_spmv_transpose_follows_value_changes(dev):
    """A^T's values are gathered on the device from A's whenever they change
    (host or device update, FP32 copy included); the pattern is kept."""
    import torch
    A = read_mtx("rdist1.mtx")
    x = la.gen.uniform(A.n, 3)
    dA = la.DeviceMatrix(as_lib_csr(A))
    xt = torch.tensor(x, device=dev)
    rng = np.random.default_rng(4)
    for step, on_device in enumerate([False, True, False]):
        av = rng.standard_normal(A.nnz)
        if on_device:
            dA.set_values(torch.tensor(av, device=dev))
        else:
            dA.set_values(av)
        At = orc.transpose(orc.CSR(A.n, A.ai, A.aj, av))
        y = torch.zeros(A.n, dtype=torch.float64, device=dev)
        dA.spmv_transpose(xt, y)
        ref = orc.spmv(At, x)
        assert np.all(np.abs(y.cpu().numpy() - ref) <= 1e-12 * _absbound(At, x) + 1e-300), step
        y32 = torch.zeros(A.n, dtype=torch.float32, device=dev)
        dA.spmv_transpose(xt.float(), y32, val=la.F32)
        assert np.all(np.abs(y32.double().cpu().numpy() - ref) <= 2e-6 * _absbound(At, x) + 1e-30)


def test_documented_spmv_deviations(dev):
    """Two documented parity exceptions, intended: (1) SegSumSPMV does not
    special-case beta == 0 (spmv.hpp:171-173), so NaN in y survives in the
    oracle's restatement; the device's SEGSUM follows SerialSPMV's rule
    (beta == 0 never reads y, spmv.hpp:26). (2) ALBUSSPMV never visits
    leading empty rows (spmv.hpp:213-218) and leaves them as they were; the
    device writes them (0 for beta == 0, beta*y otherwise) like every other
    row."""
    import torch
    D = random_sparse(200, 0.05, seed=11)
    D[:4] = 0.0
    A = from_dense(D)
    x = la.gen.uniform(A.n, 3)
    dA = la.DeviceMatrix(as_lib_csr(A))
    y_nan = np.full(A.n, np.nan)
    assert np.isnan(orc.spmv(A, x, y_nan, beta=0.0, alg="segsum", nthreads=4)).all()
    yt = torch.tensor(y_nan, device=dev)
    dA.spmv(torch.tensor(x, device=dev), yt, alg=la.SPMV_SEGSUM)
    np.testing.assert_allclose(yt.cpu().numpy(), orc.spmv(A, x), rtol=1e-12, atol=1e-14)
    y0 = la.gen.uniform(A.n, 4)
    ref_albus = orc.spmv(A, x, y0, alpha=1.0, beta=0.5, alg="albus", nthreads=8)
    np.testing.assert_array_equal(ref_albus[:4], y0[:4])        # never visited
    yt = torch.tensor(y0, device=dev)
    dA.spmv(torch.tensor(x, device=dev), yt, alpha=1.0, beta=0.5, alg=la.SPMV_ALBUS)
    np.testing.assert_allclose(yt.cpu().numpy()[:4], 0.5 * y0[:4], rtol=0, atol=0)
    np.testing.assert_allclose(yt.cpu().numpy()[4:], ref_albus[4:], rtol=1e-12, atol=1e-13)


def test_colblock_scattered_matrix(dev):
    """A ragged matrix with columns scattered over an x larger than the L2
    caches: AUTO runs the row-panel x column-block kernel (each panel's row
    sums held across the blocks); alpha/beta, the FP32/BF16 value copies and
    value updates all follow, against the oracle (FP64 1e-12 element-wise),
    and repeated products are bitwise equal."""
    import torch
    A = _o(la.gen.powerlaw(700_000, 8.0, 1.0, 3000, seed=6))
    dA = la.DeviceMatrix(as_lib_csr(A))
    assert dA.resolve(la.SPMV_AUTO) == la.SPMV_COLBLOCK
    assert dA.resolve(la.SPMV_ALBUS) == la.SPMV_ALBUS
    x = la.gen.uniform(A.n, 7)
    y0 = la.gen.uniform(A.n, 8)
    xt = torch.tensor(x, device=dev)
    for alpha, beta in [(1.0, 0.0), (2.5, -0.5), (-1.0, 1.0)]:
        yt = torch.tensor(y0, device=dev)
        dA.spmv(xt, yt, alpha=alpha, beta=beta, alg=la.SPMV_AUTO)
        ref = orc.spmv(A, x, y0, alpha=alpha, beta=beta)
        scale = abs(alpha) * _absbound(A, x) + abs(beta) * np.abs(y0)
        assert np.all(np.abs(yt.cpu().numpy() - ref) <= 1e-12 * scale + 1e-300), (alpha, beta)
    for val in (la.F32, la.BF16):
        yt = torch.zeros(A.n, dtype=torch.float32, device=dev)
        dA.spmv(xt.float(), yt, alg=la.SPMV_AUTO, val=val, acc=la.F32)
        ref_r = orc.spmv(_round_values(A, val), x)
        assert np.all(np.abs(yt.double().cpu().numpy() - ref_r) <= ACC32_BOUND * _absbound(A, x))
    av = np.random.default_rng(3).standard_normal(A.nnz)
    dA.set_values(av)
    yt = torch.zeros(A.n, dtype=torch.float64, device=dev)
    dA.spmv(xt, yt, alg=la.SPMV_AUTO)
    A2 = orc.CSR(A.n, A.ai, A.aj, av)
    assert np.all(np.abs(yt.cpu().numpy() - orc.spmv(A2, x)) <= 1e-12 * _absbound(A2, x) + 1e-300)
    for _ in range(3):
        y2 = torch.zeros_like(yt)
        dA.spmv(xt, y2, alg=la.SPMV_AUTO)
        assert torch.equal(y2, yt)


def _scattered(n, ncols, lens, seed):
    """rows of the given lengths, distinct uniform random columns in CSR order"""
    rng = np.random.default_rng(seed)
    ai = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    aj = np.empty(ai[-1], np.int64)
    for r in np.flatnonzero(lens > 1):
        aj[ai[r]:ai[r + 1]] = np.sort(rng.choice(ncols, lens[r], replace=False))
    one = np.flatnonzero(lens == 1)
    aj[ai[one]] = rng.integers(0, ncols, len(one))
    return orc.CSR(n, ai, aj, rng.standard_normal(ai[-1]), ncols=ncols)


@pytest.mark.parametrize("case", ["short_rows_full_panels", "rect_long_rows"])
def test_colblock_panel_shapes(dev, case):
    """Edge shapes of the row-panel x column-block layout: one entry per row
    over 4.3M rows with every 7th row empty (panels capped at 8192 rows, so
    more panels than workgroups fit at once), and a 60000 x 3.1M matrix whose
    rows reach 6000 entries (a row spans many lanes, several 256-entry steps
    and every block). FP64 against the oracle (1e-12 element-wise), FP32
    values against the oracle on rounded values, repeats bitwise equal."""
    import torch
    if case == "short_rows_full_panels":
        n = ncols = 4_300_000
        lens = np.ones(n, np.int64)
        lens[::7] = 0
    else:
        n, ncols = 60_000, 3_100_000
        lens = np.random.default_rng(2).integers(0, 40, n)
        lens[::997] = 6000
    A = _scattered(n, ncols, lens, 11)
    dA = la.DeviceMatrix(as_lib_csr(A))
    assert dA.resolve(la.SPMV_AUTO) == la.SPMV_COLBLOCK
    x = la.gen.uniform(ncols, 4)
    y0 = la.gen.uniform(n, 5)
    xt = torch.tensor(x, device=dev)
    yt = torch.tensor(y0, device=dev)
    dA.spmv(xt, yt, alpha=1.5, beta=-2.0, alg=la.SPMV_AUTO)
    ref = orc.spmv(A, x, y0, alpha=1.5, beta=-2.0)
    scale = 1.5 * _absbound(A, x) + 2.0 * np.abs(y0)
    got = yt.cpu().numpy()
    assert np.all(np.abs(got - ref) <= 1e-12 * scale + 1e-300)
    y2 = torch.tensor(y0, device=dev)
    dA.spmv(xt, y2, alpha=1.5, beta=-2.0, alg=la.SPMV_AUTO)
    assert torch.equal(y2, yt)
    yf = torch.zeros(n, dtype=torch.float32, device=dev)
    dA.spmv(xt.float(), yf, alg=la.SPMV_AUTO, val=la.F32, acc=la.F32)
    ref_r = orc.spmv(_round_values(A, la.F32), x)
    assert np.all(np.abs(yf.double().cpu().numpy() - ref_r) <= ACC32_BOUND * _absbound(A, x))


def test_value_grad(dev):
    import torch
    A = read_mtx("rdist1.mtx")
    lam = la.gen.uniform(A.n, 1)
    x = la.gen.uniform(A.n, 2)
    dA = la.DeviceMatrix(as_lib_csr(A))
    g = torch.zeros(A.nnz, dtype=torch.float64, device=dev)
    dA.value_grad(torch.tensor(lam, device=dev), torch.tensor(x, device=dev), g)
    rows = np.repeat(np.arange(A.n), np.diff(A.ai))
    ref = -lam[rows] * x[A.aj - A.base]
    np.testing.assert_array_equal(g.cpu().numpy(), ref)


def test_sell_follows_new_values(dev):
    """The sliced-ELL copy is re-gathered after the CSR values change, and
    a NaN in x outside a row's columns stays out of that row."""
    import torch
    A = _o(la.gen.laplace3d(20))
    dA = la.DeviceMatrix(as_lib_csr(A))
    x = la.gen.uniform(A.n, 5)
    xt = torch.tensor(x, device=dev)
    y = torch.zeros_like(xt)
    dA.spmv(xt, y, alg=la.SPMV_SELL)
    assert np.all(np.abs(y.cpu().numpy() - orc.spmv(A, x)) <= 1e-13 * _absbound(A, x))
    v2 = np.asarray(A.av) * -2.5
    dA.set_values(torch.tensor(v2, device=dev))
    dA.spmv(xt, y, alg=la.SPMV_SELL)
    A2 = orc.CSR(A.n, A.ai, A.aj, v2)
    assert np.all(np.abs(y.cpu().numpy() - orc.spmv(A2, x)) <= 1e-13 * _absbound(A2, x))
    # row 0 of the 7-point operator touches columns 0, 1, 20, 400 only
    xn = x.copy()
    xn[A.n - 1] = np.nan
    dA.spmv(torch.tensor(xn, device=dev), y, alg=la.SPMV_SELL)
    assert np.isfinite(y.cpu().numpy()[0])


def test_spmv_repeat_deterministic(dev):
    """nnz-balanced kernel with cross-segment rows is run-to-run bitwise
    identical (the fix-up adds partials in a fixed order)."""
    import torch
    A = _o(la.gen.powerlaw(20000, 16.0, 2.0, 6000, seed=5))
    x = torch.tensor(la.gen.uniform(A.n, 7), device=dev)
    dA = la.DeviceMatrix(as_lib_csr(A))
    outs = []
    for _ in range(5):
        y = torch.zeros(A.n, dtype=torch.float64, device=dev)
        dA.spmv(x, y, alg=la.SPMV_ALBUS)
        outs.append(y.cpu().numpy())
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def test_resolve_names_the_kernel(dev):
    """AUTO runs sliced ELL on a stencil (small padding) and the segment
    kernels on long, ragged rows; an empty matrix runs the row kernel."""
    st = la.DeviceMatrix(la.gen.laplace3d(12))
    assert st.resolve(la.SPMV_AUTO) == la.SPMV_SELL
    assert st.resolve(la.SPMV_SEGSUM) == la.SPMV_ALBUS
    assert st.resolve(la.SPMV_ROW) == la.SPMV_ROW
    pl = la.DeviceMatrix(la.gen.powerlaw(20000, 16.0, 2.0, 6000, seed=5))
    assert pl.resolve(la.SPMV_AUTO) == la.SPMV_ALBUS
    assert pl.resolve(la.SPMV_SELL) == la.SPMV_ALBUS
    em = la.DeviceMatrix(as_lib_csr(orc.CSR(7, np.zeros(8), [], [])))
    assert em.resolve(la.SPMV_AUTO) == la.SPMV_ROW
    # the sliced-ELL copy: 16-bit column offsets exactly when they fit
    A = la.gen.laplace3d(12)
    info = st.sell_info()
    assert info["built"] and info["col16"] and info["slots"] >= A.nnz
    assert la.DeviceMatrix(as_lib_csr(_band(70000, 32767))).sell_info()["col16"]
    wide = la.DeviceMatrix(as_lib_csr(_band(70000, 32768))).sell_info()
    assert wide["built"] and not wide["col16"]
    assert not pl.sell_info()["built"]
