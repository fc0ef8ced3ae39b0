"""Adjoint of x = A(vals)^{-1} b as a torch.autograd.Function -- a TEST
BINDING over the C ABI, not part of the hot path.

forward: the CSR values `vals` (FP64, device) are written into the device
matrix, the preconditioner is refactored on the device (IC(k) or ILU(k)
numeric phase, la_prec_refactor; the symbolic phase ran once on the host)
and the system is solved by PCG (SPD) or FGMRES through the C ABI.
backward: the adjoint system A^T lambda = dL/dx is solved with the
transposed product and the transposed preconditioner apply
(la_pcg_solve_transpose / la_fgmres_solve_transpose); then
dL/db = lambda and dL/dvals[k] = -lambda[row k] * x[col k]
(la_csr_value_grad), the gradient of -lambda^T A x on the sparsity pattern.
"""
import numpy as np
import torch

import linearalgebra_amd as la


class SparseSystem:
    """Pattern, solver settings and device handles shared by forward and
    backward. `kind` is "ic" (symmetric positive definite, PCG) or "ilu"
    (FGMRES). `precision` is one of la.PREC_*. `adjoint_rel_tol` (default
    `rel_tol`) is the relative-residual target of the backward solve: an
    adjoint right-hand side such as 2(x - x*) is smooth, and for smooth
    right-hand sides the attainable FP64 relative residual is about
    eps * cond(A), above a forward tolerance that a rough b still reaches."""

    def __init__(self, A, kind="ic", level=0, precision=la.PREC_FP64, rel_tol=1e-12,
                 max_iters=5000, trsv_mode=la.TRSV_SYNCFREE, inner_tol=1e-5,
                 adjoint_rel_tol=None):
        self.A = la.CSR(A.n, A.ai, A.aj, A.av)
        self.kind, self.level = kind, level
        self.dev = la.DeviceMatrix(self.A)
        self.opts = dict(rel_tol=rel_tol, max_iters=max_iters, trsv_mode=trsv_mode,
                         precision=precision, inner_tol=inner_tol)
        self.adjoint_rel_tol = rel_tol if adjoint_rel_tol is None else adjoint_rel_tol
        self.last = {}
        if kind == "ic":
            self.P = la.Preconditioner.ic_gpu(self.A, level)
        else:
            self.P = la.Preconditioner.ilu_gpu(self.A, level)

    def refactor(self, vals):
        """Numeric factorisation on the device from the FP64 values."""
        self.P.refactor(vals)

    def solve(self, b, transpose=False):
        x = torch.zeros_like(b)
        if self.kind == "ic":
            fn = la.pcg_transpose if transpose else la.pcg
        else:
            fn = la.fgmres_transpose if transpose else la.fgmres
        opts = dict(self.opts)
        if transpose:
            opts["rel_tol"] = self.adjoint_rel_tol
        res = fn(self.dev, self.P, b, x, raise_on_fail=True, **opts)
        self.last["adjoint" if transpose else "forward"] = res
        return x


class SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, b, system):
        v = vals.detach().contiguous()
        system.dev.set_values(v)
        system.refactor(v)
        x = system.solve(b.detach().contiguous())
        ctx.system = system
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        (x,) = ctx.saved_tensors
        S = ctx.system
        lam = S.solve(grad_x.contiguous(), transpose=True)
        gvals = torch.zeros(S.A.nnz, dtype=torch.float64, device=x.device)
        S.dev.value_grad(lam, x, gvals)
        return gvals, lam, None


def sparse_solve(vals, b, system):
    """x = A(vals)^{-1} b, differentiable in vals and b."""
    return SparseSolve.apply(vals, b, system)


def edge_values(theta, base_vals, val_map):
    """CSR values of the 2-D edge-conductivity diffusion operator as a
    differentiable function of the edge conductivities theta:
    A_ij = A_ji = -theta_e, A_ii and A_jj gain +theta_e; base_vals holds the
    theta-independent part (the Dirichlet ground terms on the diagonal)."""
    vm = val_map
    v = base_vals.clone()
    v = v.index_add(0, vm[:, 0], -theta)
    v = v.index_add(0, vm[:, 1], -theta)
    v = v.index_add(0, vm[:, 2], theta)
    v = v.index_add(0, vm[:, 3], theta)
    return v
