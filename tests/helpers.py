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


def bf16_rne(x):
    """FP64 -> the nearest BF16 value (8 significant bits, FP32's exponent
    range, subnormals k * 2^-133), ties to even, returned as float64; NaN
    stays NaN and |x| at or past the halfway point above the largest BF16
    becomes inf. Exact integer-valued arithmetic: x is scaled by the power
    of two that puts its BF16 spacing at 1, rounded by rint (half to even)
    and scaled back."""
    x = np.asarray(x, np.float64)
    out = np.array(x, copy=True)
    fin = np.isfinite(x) & (x != 0)
    xf = x[fin]
    _, e = np.frexp(xf)               # |xf| = m 2^e, 0.5 <= m < 1
    e = np.maximum(e, -125)           # the subnormal range keeps spacing 2^-133
    spacing_exp = e - 8               # 8 significant bits
    q = np.rint(np.ldexp(xf, -spacing_exp))
    r = np.ldexp(q, spacing_exp)
    r[np.abs(r) >= 2.0 ** 128] = np.inf * np.sign(r[np.abs(r) >= 2.0 ** 128])
    out[fin] = r
    return out


def bf16_bits(x):
    """BF16 bit patterns of values already on the BF16 grid (bf16_rne)."""
    return (np.asarray(x, np.float32).view(np.uint32) >> 16).astype(np.uint16)


def bf16_special_values():
    """Doubles that probe the FP64 -> BF16 rounding: exact ties (even and
    odd neighbours), one FP64 ulp either side of a tie, the subnormal range
    (ties at k + 1/2 times 2^-133, below half the smallest subnormal), the
    largest finite values and the overflow threshold, signed zeros, inf."""
    v = []
    for m in (1.0, 1.5, 1.0078125, 1.9921875, 1.25):
        for ex in (-126, -100, -3, 0, 7, 100, 127):
            base = m * 2.0 ** ex
            half = 2.0 ** (ex - 8)    # half a BF16 ulp in this binade
            for t in (base + half, base - half):
                v += [t, np.nextafter(t, np.inf), np.nextafter(t, -np.inf)]
            v.append(base)
    sub = 2.0 ** -133
    for k in (0.5, 1.0, 1.5, 2.5, 3.5, 126.5, 127.5, 127.75, 0.25, 0.75):
        t = k * sub
        v += [t, np.nextafter(t, np.inf), np.nextafter(t, -np.inf)]
    v += [2.0 ** -149, 2.0 ** -1074, 2.0 ** -126 * (1 - 2.0 ** -20)]
    mx = (2 - 2.0 ** -7) * 2.0 ** 127  # largest BF16
    thr = 2.0 ** 128 - 2.0 ** 119      # halfway to 2^128: rounds to even (inf)
    v += [mx, np.nextafter(thr, 0), thr, np.nextafter(thr, np.inf), 2.0 ** 128, 1e300,
          0.0, -0.0, np.inf]
    v = np.array(v, np.float64)
    return np.concatenate([v, -v])


# the lower-row offsets (di, dj, dk) of a 27-point stencil's ILU(1) factor
_ILU1_27 = [(di, dj, -1) for dj in range(-1, 3) for di in range(-2, 3)] + \
    [(di, dj, 0) for dj in (-2, -1) for di in range(-2, 3)] + [(-2, 0, 0), (-1, 0, 0)]


def random_box_triangle(nx, ny, nz, seed, upper):
    """A strict triangle on a natural-order box grid whose rows hold a
    random subset of the 27-point ILU(1) lower-row offsets (always the
    neighbours one step back along each axis), with random values."""
    rng = np.random.default_rng(seed)
    must = [(-1, 0, 0), (0, -1, 0), (0, 0, -1)]
    rest = [o for o in _ILU1_27 if o not in must]
    rows, cols = [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = i + nx * (j + ny * k)
                cs = set()
                for di, dj, dk in must + [o for o in rest if rng.random() < 0.6]:
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                        cs.add(ii + nx * (jj + ny * kk))
                for c in sorted(cs):
                    rows.append(r)
                    cols.append(c)
    n = nx * ny * nz
    rows, cols = np.array(rows), np.array(cols)
    if upper:  # mirror the pattern: row n-1-r reads n-1-c
        rows, cols = n - 1 - rows, n - 1 - cols
        o = np.lexsort((cols, rows))
        rows, cols = rows[o], cols[o]
    ai = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    av = rng.uniform(-1, 1, len(cols)) / 64  # rows of at most 31: diagonally dominant
    return orc.CSR(n, ai, cols.astype(np.int32), av)
