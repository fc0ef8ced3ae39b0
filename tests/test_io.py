"""Readers (la_read_matrix_market, la_read_binary_coo, la_read_binary_vector)
against independent restatements: scipy's Matrix Market reader + a
(row, col) sort for .mtx files; a numpy decoding of the 16-byte records
(value at byte 0, column at 8, row at 12 -- std::tuple<int32,int32,double>
in libstdc++) + a (row, col) sort for the reference's binary COO files
(tests/golden/{K,M,G,G2}.bin, copied from the reference's data/shared/).
Matrix Market symmetric expansion stores a diagonal entry once; the
reference's reader (fast_matrix_market, unpinned version) is not available,
so that convention is parity unpinned beyond scipy's."""
import os

import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import GOLDEN, read_mtx


@pytest.mark.parametrize("name", ["ex5.mtx", "rdist1.mtx", "jgl009.mtx"])
@pytest.mark.parametrize("base", [0, 1])
def test_matrix_market_matches_scipy(name, base):
    A = la.read_matrix_market(os.path.join(GOLDEN, name), base=base)
    R = read_mtx(name, base=base)
    assert (A.n, A.ncols) == (R.n, R.ncols)
    assert np.array_equal(A.ai, R.ai) and np.array_equal(A.aj, R.aj)
    assert np.array_equal(A.av, R.av)


def test_ex5_sizes():
    # 27 x 27 symmetric, 153 stored entries -> 279 after expansion
    A = la.read_matrix_market(os.path.join(GOLDEN, "ex5.mtx"))
    assert (A.n, A.nnz) == (27, 279)


def _dense_of(A):
    D = np.zeros((A.n, A.ncols))
    b = A.base
    for i in range(A.n):
        for q in range(A.ai[i] - b, A.ai[i + 1] - b):
            D[i, A.aj[q] - b] += A.av[q]
    return D


def test_symmetry_kinds(tmp_path):
    entries = [(1, 1, 2.0), (2, 1, -1.5), (3, 2, 4.0), (3, 3, 1.0)]
    body = "".join("%d %d %r\n" % e for e in entries)
    for sym, sgn in [("general", None), ("symmetric", 1.0), ("skew-symmetric", -1.0)]:
        f = tmp_path / ("m_%s.mtx" % sym)
        f.write_text("%%%%MatrixMarket matrix coordinate real %s\n%% comment\n3 3 4\n%s" % (sym, body))
        A = la.read_matrix_market(str(f))
        D = np.zeros((3, 3))
        for i, j, v in entries:
            D[i - 1, j - 1] = v
            if sgn is not None and i != j:
                D[j - 1, i - 1] = sgn * v
        assert np.array_equal(_dense_of(A), D), sym
    f = tmp_path / "p.mtx"
    f.write_text("%%MatrixMarket matrix coordinate pattern general\n2 3 2\n1 3\n2 1\n")
    A = la.read_matrix_market(str(f))
    assert (A.n, A.ncols) == (2, 3) and list(A.av) == [1.0, 1.0]


def test_matrix_market_rejects_bad_files(tmp_path):
    f = tmp_path / "bad.mtx"
    f.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(la.LaError):
        la.read_matrix_market(str(f))
    with pytest.raises(la.LaError):
        la.read_matrix_market(str(tmp_path / "missing.mtx"))


def _numpy_bin(path, base):
    b = open(path, "rb").read()
    n = int(np.frombuffer(b[:8], np.int64)[0])
    rec = np.frombuffer(b[8:8 + 16 * n], dtype=[("v", "<f8"), ("c", "<i4"), ("r", "<i4")])
    r, c, v = rec["r"].astype(np.int64), rec["c"].astype(np.int64), rec["v"]
    rows, cols = int(r.max()) + 1 - base, int(c.max()) + 1 - base
    o = np.lexsort((c, r))
    ai = np.zeros(rows + 1, np.int64)
    np.add.at(ai, r[o] - base + 1, 1)
    return rows, cols, np.cumsum(ai) + base, c[o], v[o]


@pytest.mark.parametrize("name,shape", [("K.bin", (125, 125)), ("M.bin", (125, 125)),
                                        ("G.bin", (37, 2)), ("G2.bin", (42577, 52))])
def test_binary_coo_matches_numpy(name, shape):
    path = os.path.join(GOLDEN, name)
    A = la.read_binary_coo(path, base=1)
    rows, cols, ai, aj, av = _numpy_bin(path, 1)
    assert (A.n, A.ncols) == (rows, cols) == shape
    assert np.array_equal(A.ai, ai) and np.array_equal(A.aj, aj) and np.array_equal(A.av, av)


def test_binary_stiffness_is_symmetric():
    for name in ("K.bin", "M.bin"):
        D = _dense_of(la.read_binary_coo(os.path.join(GOLDEN, name), base=1))
        assert np.allclose(D, D.T, rtol=0, atol=1e-12 * np.abs(D).max())


def test_binary_vector(tmp_path):
    v = np.random.default_rng(3).standard_normal(17)
    f = tmp_path / "v.bin"
    f.write_bytes(np.int64(17).tobytes() + v.tobytes())
    assert np.array_equal(la.read_binary_vector(str(f)), v)
