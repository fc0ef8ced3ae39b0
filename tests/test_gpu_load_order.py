"""Regression: the library loaded (by a host-only call) before torch has
touched the GPU must still reach the device. torch ships its own HIP
runtime; both runtimes share one HSA runtime in the process, and the
library's device calls failed with 'no ROCm-capable device' when it came up
first. Runs in a fresh interpreter so the import order is the tested one."""
import os
import subprocess
import sys

import pytest

from helpers import ROOT

pytestmark = pytest.mark.gpu

SCRIPT = r"""
import sys
sys.path.insert(0, ROOT)
import numpy as np
import linearalgebra_amd as la
A = la.gen.poisson2d(16)            # host-only call: loads the library first
import torch
x = torch.ones(A.n, dtype=torch.float64, device="cuda:0")
y = torch.zeros_like(x)
la.DeviceMatrix(A).spmv(x, y)
rows = np.diff(A.ai)
assert float(y.sum()) == float(np.sum(A.av)), (float(y.sum()), float(np.sum(A.av)))
print("ok")
"""


def test_library_first_then_torch(dev):
    env = dict(os.environ)
    r = subprocess.run([sys.executable, "-c", SCRIPT.replace("ROOT", repr(ROOT))],
                       capture_output=True, text=True, timeout=110, env=env)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "ok" in r.stdout
