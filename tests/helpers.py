"""Shared test helpers: fixture loading and small matrix builders."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402  (CPU oracle)


def read_mtx(name, base=0):
    """Matrix Market -> CSR with rows/columns sorted (utils/utils.h:93-132:
    triplets sorted by (row, col), then row counts prefix-summed). Symmetric
    files are expanded to both triangles."""
    import scipy.io
    M = scipy.io.mmread(os.path.join(GOLDEN, name)).tocoo()
    r, c, v = M.row.astype(np.int64), M.col.astype(np.int64), M.data.astype(np.float64)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    n = M.shape[0]
    ai = np.zeros(n + 1, np.int64)
    np.add.at(ai, r + 1, 1)
    ai = np.cumsum(ai)
    return orc.CSR(n, ai + base, c + base, v, ncols=M.shape[1])


def from_dense(D, base=0):
    D = np.asarray(D, dtype=np.float64)
    n = D.shape[0]
    ai, aj, av = [base], [], []
    for i in range(n):
        for j in range(D.shape[1]):
            if D[i, j] != 0:
                aj.append(j + base)
                av.append(D[i, j])
        ai.append(len(aj) + base)
    return orc.CSR(n, ai, aj, av, ncols=D.shape[1])


def to_dense(A):
    b = A.base
    D = np.zeros((A.n, A.ncols))
    for i in range(A.n):
        for k in range(A.ai[i] - b, A.ai[i + 1] - b):
            D[i, A.aj[k] - b] += A.av[k]
    return D


def as_lib_csr(A):
    import linearalgebra_amd as la
    return la.CSR(A.n, A.ai, A.aj, A.av, ncols=A.ncols)


def as_orc_csr(A):
    return orc.CSR(A.n, A.ai, A.aj, A.av, ncols=A.ncols)


def random_sparse(n, density, seed, diag=None, ncols=None):
    rng = np.random.default_rng(seed)
    m = n if ncols is None else ncols
    D = rng.standard_normal((n, m)) * (rng.random((n, m)) < density)
    if diag is not None:
        for i in range(min(n, m)):
            D[i, i] = diag + abs(D[i]).sum()
    return D


def rel_err(x, ref):
    x, ref = np.asarray(x, np.float64), np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / (den if den > 0 else 1.0)
