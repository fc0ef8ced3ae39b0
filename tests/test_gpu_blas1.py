"""Fused BLAS-1 entry points of the C ABI (la_dot, la_axpy_dot): FP64
accumulation of the products, for FP64 and FP32 vectors, 16-byte aligned or
not, against math.fsum of the same products (exactly rounded sums); the
error bound is the one of any FP64 summation order, n eps sum |x_i y_i|."""
import ctypes as C
import math

import numpy as np
import pytest

import linearalgebra_amd as la
from linearalgebra_amd import _lib

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _dot(n, code, x, y):
    r = C.c_double()
    st = la.lib().la_dot(n, code, x, y, C.byref(r), None)
    return st, r.value


def _exact(x, y):
    p = np.asarray(x, np.float64) * np.asarray(y, np.float64)
    return math.fsum(p.tolist()), float(np.abs(p).sum())


@pytest.mark.parametrize("dt,code", [("float64", _lib.F64), ("float32", _lib.F32)])
@pytest.mark.parametrize("n,offset", [(1_000_003, 0), (1_000_003, 1), (4, 0), (3, 1), (0, 0)])
def test_dot_matches_exact_sum(dev, dt, code, n, offset):
    import torch
    rng = np.random.default_rng(n + offset)
    tdt = getattr(torch, dt)
    xs = torch.tensor(rng.standard_normal(n + offset), dtype=tdt, device=dev)
    ys = torch.tensor(rng.standard_normal(n + offset), dtype=tdt, device=dev)
    x, y = xs[offset:], ys[offset:]  # offset 1: not 16-byte aligned
    st, v = _dot(n, code, x.data_ptr() if n else None, y.data_ptr() if n else None)
    assert st == 0
    ref, mag = _exact(x.cpu().numpy(), y.cpu().numpy())
    assert abs(v - ref) <= max(n, 1) * EPS * mag


@pytest.mark.parametrize("dt,code", [("float64", _lib.F64), ("float32", _lib.F32)])
@pytest.mark.parametrize("with_z", [False, True])
def test_axpy_dot(dev, dt, code, with_z):
    import torch
    n, a = 300_001, -0.75
    rng = np.random.default_rng(7)
    tdt = getattr(torch, dt)
    x = torch.tensor(rng.standard_normal(n), dtype=tdt, device=dev)
    y = torch.tensor(rng.standard_normal(n), dtype=tdt, device=dev)
    z = torch.tensor(rng.standard_normal(n), dtype=tdt, device=dev)
    y0 = y.cpu().numpy().astype(np.float64)
    r = C.c_double()
    st = la.lib().la_axpy_dot(n, code, a, x.data_ptr(), y.data_ptr(),
                              z.data_ptr() if with_z else None, C.byref(r), None)
    assert st == 0
    y1 = y.cpu().numpy()
    # y = a x + y, computed in FP64 and rounded once to the vector type
    want = a * x.cpu().numpy().astype(np.float64) + y0
    ulp = np.spacing(np.abs(want).astype(y1.dtype)).astype(np.float64)
    tol = 2 * ulp + 4 * EPS * (np.abs(a * x.cpu().numpy()) + np.abs(y0))
    assert np.all(np.abs(y1 - want) <= tol)
    ref, mag = _exact(y1, z.cpu().numpy() if with_z else y1)
    assert abs(r.value - ref) <= n * EPS * mag
