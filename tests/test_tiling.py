"""The tiled schedule's layout (linearalgebra_amd/csrc/tiling.cpp), replayed
on the CPU (tests/cpp/tiling_capi.cpp, built here with g++; no GPU).

The replay follows the kernel's rules: tiles in ticket order, each tile's
groups in order; a group first copies its halo values (every one must
already be solved by an earlier tile: the schedule drains whatever number
of tiles is resident), then solves its rows from the tile's own window of
the last 16 groups, the halo and the zero slot, and writes its results
only after all of its reads. The solution must equal plain substitution
(triangle_solve.hpp:32-71) to 1e-12; a pattern the layout refuses keeps
the level-chunk schedule, and the reason is checked.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import ROOT, orc, random_box_triangle

LANES, WINDOW = 64, 16
ZERO_SLOT = WINDOW * LANES


@pytest.fixture(scope="module")
def tl(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    so = str(tmp_path_factory.mktemp("tiling") / "libtiling.so")
    csrc = os.path.join(ROOT, "linearalgebra_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fopenmp", "-shared", "-fPIC", "-I", csrc,
                    "-o", so, os.path.join(csrc, "tiling.cpp"),
                    os.path.join(ROOT, "tests", "cpp", "tiling_capi.cpp")],
                   check=True, timeout=300)
    lib = C.CDLL(so)
    lib.tiling_why.restype = C.c_char_p
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _levels(n, ai, aj, upper):
    lv = np.zeros(n, np.int32)
    for r in (range(n - 1, -1, -1) if upper else range(n)):
        cs = aj[ai[r]:ai[r + 1]]
        if len(cs):
            lv[r] = lv[cs].max() + 1
    return lv


def _build(tl, T, upper):
    n = T.n
    ai = np.ascontiguousarray(np.asarray(T.ai) - T.ai[0], np.int32)
    aj = np.ascontiguousarray(np.asarray(T.aj) - T.ai[0], np.int32)
    lv = _levels(n, ai, aj, upper)
    ok = tl.tiling_build(n, _p(ai), _p(aj), _p(lv), LANES, 256, n, 19000)
    return ok, tl.tiling_why().decode(), ai, aj


def _replay(tl, T, diag, b, upper):
    """x from the layout, with the kernel's read/write rules; asserts that
    every value is solved before it is read."""
    ok, why, ai, aj = _build(tl, T, upper)
    assert ok, why
    v = np.zeros(7, np.int32)
    tl.tiling_sizes(_p(v))
    nparts, ngroups, rl, E, hmax, nrf, ncode = (int(t) for t in v)
    n = T.n
    pg = np.zeros(nparts + 1, np.int32)
    gr = np.zeros(ngroups + 1, np.int32)
    tp = np.zeros(n, np.int32)
    tc = np.zeros(ncode, np.int32)
    ts = np.zeros(ncode, np.int32)
    rp = np.zeros(ngroups + 1, np.int32)
    rc = np.zeros(max(nrf, 1), np.int32)
    rs = np.zeros(max(nrf, 1), np.int32)
    tl.tiling_get(_p(pg), _p(gr), _p(tp), _p(tc), _p(ts), _p(rp), _p(rc), _p(rs))
    assert ncode == n * rl * E
    assert sorted(tp.tolist()) == list(range(n))
    av = np.asarray(T.av, np.float64)
    x = np.full(n, np.nan)
    for pi in range(nparts):
        win = np.full(ZERO_SLOT + 1, np.nan)
        win[ZERO_SLOT] = 0.0
        halo = np.full(max(hmax, 1), np.nan)
        for g in range(pg[pi], pg[pi + 1]):
            gi = g - pg[pi]
            for t in range(rp[g], rp[g + 1]):
                assert not np.isnan(x[rc[t]]), ("halo value not solved yet", pi, g)
                halo[rs[t]] = x[rc[t]]
            rows = range(gr[g], gr[g + 1])
            assert 0 < len(rows) * rl <= LANES
            outs = []
            for kk in rows:
                r = tp[kk]
                acc = 0.0
                for s in range(rl * E):
                    c, q = tc[kk * rl * E + s], ts[kk * rl * E + s]
                    val = win[c] if c >= 0 else halo[-c - 1]
                    assert not np.isnan(val), ("read before written", pi, g, kk)
                    if q >= 0:
                        assert ai[r] <= q < ai[r + 1]
                        acc += av[q] * val
                    else:
                        assert c == ZERO_SLOT
                outs.append((kk, r, (b[r] - acc) / (diag[r] if diag is not None else 1.0)))
            for kk, r, out in outs:
                base = (gi % WINDOW) * LANES + rl * (kk - gr[g])
                win[base:base + rl] = out
                x[r] = out
    return x, nparts, rl


def _o(A):
    return orc.CSR(A.n, A.ai, A.aj, A.av, ncols=A.ncols)


def _rcm(A):
    ip, p = la.rcm(A)
    return la.permute(A, ip, p)


def _cases():
    """(name, factor maker, expected lanes per row)"""
    return [
        ("poisson2d_48_ilu0", lambda: orc.iluk(_o(la.gen.poisson2d(48)), 0)[0], 1),
        ("laplace3d_12_rcm_ilu0", lambda: orc.iluk(_o(_rcm(la.gen.laplace3d(12))), 0)[0], 1),
        ("laplace3d_12x10x9_ilu0", lambda: orc.iluk(_o(la.gen.laplace3d(12, 10, 9)), 0)[0], 1),
        ("aniso27_14x12x10_ilu1", lambda: orc.iluk(_o(la.gen.aniso27(14, 12, 10)), 1)[0], 4),
        ("aniso27_12x10x9_ilu0", lambda: orc.iluk(_o(la.gen.aniso27(12, 10, 9)), 0)[0], 4),
        ("laplace3d_12x10x9_ilu2", lambda: orc.iluk(_o(la.gen.laplace3d(12, 10, 9)), 2)[0], 4),
    ]


@pytest.mark.parametrize("name,make,rl", _cases(), ids=[c[0] for c in _cases()])
def test_tiled_layout_replays_to_substitution(tl, name, make, rl):
    F = make()
    L, D, U = orc.split_ldu(F)
    b = la.gen.uniform(F.n, 21)
    for T, diag, upper, ref in ((L, None, False, orc.forward_sub(L, b)),
                                (U, D, True, orc.backward_sub(U, D, b))):
        x, nparts, lanes = _replay(tl, T, diag, b, upper)
        assert nparts > 1 and lanes == rl
        assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max(), name


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("upper", [False, True])
def test_tiled_layout_random_box_patterns(tl, seed, upper):
    """Box-grid triangles whose rows hold random subsets of the 27-point
    ILU(1) pattern get a layout that drains in ticket order and solves
    exactly, whichever entries each row holds."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    T = random_box_triangle(13, 11, 9, seed, upper)
    D = 1.0 + np.random.default_rng(seed).random(T.n)
    b = la.gen.uniform(T.n, 22)
    x, nparts, rl = _replay(tl, T, D, b, upper)
    M = sp.csr_matrix((T.av, T.aj, T.ai), shape=(T.n, T.n)) + sp.diags(D)
    ref = spl.spsolve_triangular(M.tocsr(), b, lower=not upper)
    assert nparts > 1 and rl == 4
    assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()


def test_tiled_layout_entries_two_planes_back(tl):
    """Rows of the 27-point stencil's lower half plus entries two planes
    back, (i - 2, j - 2, k - 2) and (i + 2, j + 2, k - 2) (reads from two
    k blocks back, all through the halo), replay exactly."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    nx, ny, nz = 13, 11, 9
    n = nx * ny * nz
    offs = [(di, dj, dk) for dk in (-1, 0) for dj in (-1, 0, 1) for di in (-1, 0, 1)
            if (dk, dj, di) < (0, 0, 0)] + [(-2, -2, -2), (2, 2, -2)]
    rows, cols = [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = i + nx * (j + ny * k)
                for c in sorted({(i + di) + nx * ((j + dj) + ny * (k + dk)) for di, dj, dk in offs
                                 if 0 <= i + di < nx and 0 <= j + dj < ny and 0 <= k + dk < nz}):
                    rows.append(r)
                    cols.append(c)
    ai = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    av = np.random.default_rng(5).uniform(-1, 1, len(cols)) / 32
    T = orc.CSR(n, ai, np.array(cols, np.int32), av)
    D = np.full(n, 2.0)
    b = la.gen.uniform(n, 23)
    x, nparts, rl = _replay(tl, T, D, b, False)
    M = sp.csr_matrix((T.av, T.aj, T.ai), shape=(n, n)) + sp.diags(D)
    ref = spl.spsolve_triangular(M.tocsr(), b, lower=True)
    assert nparts > 1 and rl == 4
    assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()


def test_tiled_layout_refusals(tl):
    """Patterns the tiled schedule does not take, with the reason given."""
    F = orc.iluk(_o(la.gen.aniso27(12, 10, 9)), 2)[0]  # rows of up to 57 entries
    L, _, _ = orc.split_ldu(F)
    ok, why, _, _ = _build(tl, L, False)
    assert not ok and "longer" in why
    F = orc.iluk(_o(la.gen.powerlaw(3000, 8.0, 1.0, 400, seed=2)), 0)[0]
    L, _, _ = orc.split_ldu(F)
    ok, why, _, _ = _build(tl, L, False)
    assert not ok and why
    S = orc.iluk(_o(la.gen.poisson2d(8)), 0)[0]  # 64 rows: too few
    L, _, _ = orc.split_ldu(S)
    ok, why, _, _ = _build(tl, L, False)
    assert not ok and "few" in why
