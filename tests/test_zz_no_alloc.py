"""la_dot_noalloc (the allocation-free form of la_dot, partial sums in a
fixed per-GPU device array) against la_dot, through the C ABI: bitwise
equal on two calls in a row, the second on a smaller input, in FP64 and
FP32, on 16-byte aligned and unaligned vectors."""
import ctypes as C

import numpy as np
import pytest

import linearalgebra_amd as la
from linearalgebra_amd import _lib

pytestmark = pytest.mark.gpu


def _ptr(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("offset", [0, 1])
def test_dot_noalloc_matches_dot(dev, dtype, offset):
    import torch
    tdt = getattr(torch, dtype)
    vec = _lib.F64 if dtype == "float64" else _lib.F32
    L = la.lib()
    out = torch.full((4,), float("nan"), dtype=torch.float64, device=dev)
    for n in (3_000_001, 1_000):  # the second call is smaller
        print("dot_noalloc", dtype, "offset", offset, "n", n, flush=True)
        g = torch.Generator(device="cpu").manual_seed(n)
        xb = torch.rand(n + offset, generator=g, dtype=torch.float64).to(dev, tdt) - 0.5
        yb = torch.rand(n + offset, generator=g, dtype=torch.float64).to(dev, tdt) - 0.5
        x, y = xb[offset:], yb[offset:]
        ref = C.c_double()
        assert L.la_dot(n, vec, _ptr(x), _ptr(y), C.byref(ref), None) == 0
        res = out[1:2]
        assert L.la_dot_noalloc(n, vec, _ptr(x), _ptr(y), _ptr(res), None) == 0
        torch.cuda.synchronize()
        got = float(res.item())
        assert got == ref.value, (n, got, ref.value)
        exact = float(np.dot(x.double().cpu().numpy(), y.double().cpu().numpy()))
        assert abs(got - exact) <= 1e-12 * float((x.double().abs() * y.double().abs()).sum())
    # neighbours of the result were not written
    assert torch.isnan(out[0]) and torch.isnan(out[2:]).all()


def test_dot_noalloc_rejects_bad_arguments(dev):
    import torch
    L = la.lib()
    x = torch.ones(8, dtype=torch.float64, device=dev)
    assert L.la_dot_noalloc(8, _lib.F64, _ptr(x), _ptr(x), None, None) == _lib.LA_ERR_ARG
    assert L.la_dot_noalloc(-1, _lib.F64, _ptr(x), _ptr(x), _ptr(x), None) == _lib.LA_ERR_ARG
    assert L.la_dot_noalloc(8, _lib.F64, None, _ptr(x), _ptr(x), None) == _lib.LA_ERR_ARG
    assert L.la_dot_noalloc(0, _lib.F64, None, None, _ptr(x[:1]), None) == 0
    torch.cuda.synchronize()
    assert x[0].item() == 0.0
