"""Writes c2_ic0_pcg.json: the oracle's IC(0)-PCG on the RCM-ordered 128^3
7-point Laplacian, b ~ U[-1, 1) (seed 2024), to a 1e-10 true relative
residual -- iteration count, final residual, ||x|| and 64 seeded entries of
x. The full-size GPU test compares against it without rerunning the ~20 s
serial solve. Run from the repository root: python tests/golden/make_c2_pcg.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import linearalgebra_amd as la  # noqa: E402  (host generators, RCM, IC(0))
from oracle import oracle as orc  # noqa: E402


def main():
    A = la.gen.laplace3d(128)
    ip, p = la.rcm(A)
    A = la.permute(A, ip, p)
    U = la.split_triangle(A, upper=True)
    R, ok = la.icc(U, 0)
    assert ok
    b = la.gen.uniform(A.n, 2024)
    x, conv, it, rr = orc.pcg(orc.CSR(A.n, A.ai, A.aj, A.av), b, kind="ic",
                              F=orc.CSR(R.n, R.ai, R.aj, R.av), rel_tol=1e-10, maxiter=5000)
    assert conv
    idx = np.random.default_rng(7).choice(A.n, size=64, replace=False)
    out = {"workload": "laplace3d 128^3, RCM, IC(0), b = uniform(seed 2024), rel_tol 1e-10",
           "iters": int(it), "rel_residual": float(rr), "x_norm": float(np.linalg.norm(x)),
           "sample_index": [int(i) for i in idx], "sample_x": [float(x[i]) for i in idx]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_ic0_pcg.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path, it, rr)


if __name__ == "__main__":
    main()
