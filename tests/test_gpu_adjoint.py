"""Adjoint of x = A(theta)^{-1} b for the 2-D 5-point edge-conductivity
diffusion operator (the C5 workload): loss ||x - x*||^2 with x* = A(theta*)^{-1} b.

dL/dtheta and dL/db come from the torch.autograd.Function test binding
(linearalgebra_amd/autograd.py) over the C ABI: a transposed IC(0)-PCG
adjoint solve and the value-gradient kernel. Checked against
  * an FP64 CPU reference gradient (sparse direct LU solves, scipy's
    SuperLU, + the closed form dL/dtheta_e = -(lambda_i - lambda_j)(x_i - x_j)),
    rel. error <= 1e-6; the GPU forward solves stop at a 1e-11 and the
    adjoint solves at a 1e-10 true relative residual (the adjoint
    right-hand side is smooth, and a smooth right-hand side's attainable
    FP64 relative residual for this operator is about 1e-11);
  * central finite differences on 64 seeded theta entries at 128 x 128 and
    on 16 seeded entries at the full 512 x 512 size, error <= 1e-4 of the
    largest checked component (h = 3e-4 theta_e, truncation O(h^2));
  * the mixed FP32 path against the FP64 path, rel. error <= 1e-6.
The reference-gradient check runs at the full 512 x 512 size.
"""
import numpy as np
import pytest

import linearalgebra_amd as la
from helpers import as_lib_csr, orc, rel_err

pytestmark = pytest.mark.gpu


def _setup(nx, dev, seed=5):
    import torch
    from linearalgebra_amd.autograd import SparseSystem, edge_values
    ne = 2 * nx * nx - 2 * nx
    theta = la.gen.lognormal(ne, seed, 0.0, 0.5)
    theta_star = la.gen.lognormal(ne, seed + 50, 0.0, 0.5)
    A0, ei, ej, vm = la.gen.diffusion2d(nx, nx, np.zeros(ne))
    b = la.gen.uniform(nx * nx, seed + 500)
    vm_t = torch.tensor(vm, dtype=torch.int64, device=dev)
    base = torch.tensor(A0.av, device=dev)
    return dict(theta=theta, theta_star=theta_star, A0=A0, ei=ei, ej=ej, vm=vm, vm_t=vm_t,
                base=base, b=b, edge_values=edge_values, SparseSystem=SparseSystem)


def _loss_and_grads(S, th, th_star, system, dev):
    import torch
    from linearalgebra_amd.autograd import sparse_solve
    b = torch.tensor(S["b"], device=dev, requires_grad=True)
    theta = torch.tensor(th, device=dev, requires_grad=True)
    with torch.no_grad():
        xs = sparse_solve(S["edge_values"](torch.tensor(th_star, device=dev), S["base"], S["vm_t"]),
                          b.detach(), system)
    x = sparse_solve(S["edge_values"](theta, S["base"], S["vm_t"]), b, system)
    L = ((x - xs) ** 2).sum()
    L.backward()
    return float(L), theta.grad.cpu().numpy(), b.grad.cpu().numpy(), x.detach().cpu().numpy()


def _reference_gradient(S, nx):
    """FP64 CPU: direct sparse LU solves and the closed form."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    A, _, _, _ = la.gen.diffusion2d(nx, nx, S["theta"])
    As, _, _, _ = la.gen.diffusion2d(nx, nx, S["theta_star"])
    def lu(M):
        return spl.splu(sp.csr_matrix((M.av, M.aj, M.ai), shape=(M.n, M.n)).tocsc(),
                        permc_spec="MMD_AT_PLUS_A")
    LA = lu(A)
    x = LA.solve(S["b"])
    xs = lu(As).solve(S["b"])
    lam = LA.solve(2 * (x - xs))  # A symmetric: A^T = A
    g = -(lam[S["ei"]] - lam[S["ej"]]) * (x[S["ei"]] - x[S["ej"]])
    return g, lam, x


def as_orc(A):
    return orc.CSR(A.n, A.ai, A.aj, A.av)


def test_adjoint_vs_fp64_reference_full_size(dev):
    nx = 512
    S = _setup(nx, dev)
    system = S["SparseSystem"](S["A0"], kind="ic", level=0, rel_tol=1e-11, max_iters=20000,
                               adjoint_rel_tol=1e-10)
    print("C5 512x512 adjoint", flush=True)
    L, g, gb, x = _loss_and_grads(S, S["theta"], S["theta_star"], system, dev)
    print(system.last, flush=True)
    g_ref, lam_ref, x_ref = _reference_gradient(S, nx)
    assert rel_err(x, x_ref) < 1e-8
    assert rel_err(g, g_ref) < 1e-6, rel_err(g, g_ref)
    assert rel_err(gb, lam_ref) < 1e-6


@pytest.mark.parametrize("nx,count", [(128, 64), (512, 16)])
def test_adjoint_finite_differences(dev, nx, count):
    S = _setup(nx, dev, seed=9)
    system = S["SparseSystem"](S["A0"], kind="ic", level=0, rel_tol=1e-11, max_iters=20000,
                               adjoint_rel_tol=1e-10)
    L0, g, gb, _ = _loss_and_grads(S, S["theta"], S["theta_star"], system, dev)
    rng = np.random.default_rng(17)
    idx = rng.choice(len(S["theta"]), size=count, replace=False)
    fd = np.empty(count)
    if nx > 128:  # finer grid, larger condition number: tighter loss solves
        system = S["SparseSystem"](S["A0"], kind="ic", level=0, rel_tol=1e-12, max_iters=40000)
    for t, e in enumerate(idx):
        h = 3e-4 * S["theta"][e]
        tp, tm = S["theta"].copy(), S["theta"].copy()
        tp[e] += h
        tm[e] -= h
        Lp = _loss_only(S, tp, system, dev)
        Lm = _loss_only(S, tm, system, dev)
        fd[t] = (Lp - Lm) / (2 * h)
    err = np.abs(fd - g[idx]) / (np.abs(g[idx]) + 1e-12 * np.abs(g).max())
    print("nx", nx, "fd rel err max", err.max(), "median", np.median(err),
          "max abs err / max |g|", np.max(np.abs(fd - g[idx])) / np.abs(g[idx]).max(), flush=True)
    assert np.max(np.abs(fd - g[idx])) <= 1e-4 * np.abs(g[idx]).max()
    assert np.median(err) < 1e-4


def _loss_only(S, th, system, dev):
    import torch
    from linearalgebra_amd.autograd import sparse_solve
    with torch.no_grad():
        b = torch.tensor(S["b"], device=dev)
        if "_xs" not in S:
            S["_xs"] = sparse_solve(S["edge_values"](torch.tensor(S["theta_star"], device=dev),
                                                     S["base"], S["vm_t"]), b, system)
        x = sparse_solve(S["edge_values"](torch.tensor(th, device=dev), S["base"], S["vm_t"]),
                         b, system)
        return float(((x - S["_xs"]) ** 2).sum())


def test_adjoint_mixed_precision_matches_fp64(dev):
    nx = 192
    S = _setup(nx, dev, seed=3)
    kw = dict(kind="ic", level=0, rel_tol=1e-11, max_iters=20000, adjoint_rel_tol=1e-10)
    sys64 = S["SparseSystem"](S["A0"], **kw)
    sys32 = S["SparseSystem"](S["A0"], precision=la.PREC_MIXED_F32, inner_tol=1e-5, **kw)
    _, g64, gb64, _ = _loss_and_grads(S, S["theta"], S["theta_star"], sys64, dev)
    _, g32, gb32, _ = _loss_and_grads(S, S["theta"], S["theta_star"], sys32, dev)
    print(sys32.last, flush=True)
    assert rel_err(g32, g64) < 1e-6
    assert rel_err(gb32, gb64) < 1e-6


def test_transposed_solve_nonsymmetric_ilu(dev):
    """FGMRES + ILU(0) on a non-symmetric matrix: the adjoint solve
    A^T lambda = g (transposed product and M^-T apply) against a dense solve,
    and the value gradient against its definition."""
    import torch
    from linearalgebra_amd.autograd import SparseSystem, sparse_solve
    A = la.gen.powerlaw(1500, 8.0, 1.0, 200, seed=8)
    system = SparseSystem(A, kind="ilu", level=0, rel_tol=1e-13, max_iters=2000)
    vals = torch.tensor(A.av, device=dev, requires_grad=True)
    b = torch.tensor(la.gen.uniform(A.n, 1), device=dev, requires_grad=True)
    w = torch.tensor(la.gen.uniform(A.n, 2), device=dev)
    x = sparse_solve(vals, b, system)
    (x * w).sum().backward()
    D = as_orc(A).to_scipy().toarray()
    lam = np.linalg.solve(D.T, w.cpu().numpy())
    xr = np.linalg.solve(D, b.detach().cpu().numpy())
    assert rel_err(b.grad.cpu().numpy(), lam) < 1e-10
    rows = np.repeat(np.arange(A.n), np.diff(A.ai))
    assert rel_err(vals.grad.cpu().numpy(), -lam[rows] * xr[A.aj]) < 1e-10
