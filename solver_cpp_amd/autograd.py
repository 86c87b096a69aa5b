"""Differentiable sparse solve x = A(vals)^{-1} b as a torch.autograd.Function.

Forward: ks_solve on the GPU (any solver the C ABI accepts, e.g. mixed
FP32-inner VPCG / VPGMRES); the solution stays resident on the device and is
saved for the backward pass.

Backward (hand-written adjoint, no reference counterpart): for a loss L(x),
  A^T lambda = dL/dx          (ks_adjoint: same Krylov family on A^T - CG
                               for SPD matrices, GMRES/BiCG on the explicit
                               transposed CSR otherwise)
  dL/db      = lambda
  dL/dA_ij   = -lambda_i x_j   on the sparsity pattern (SDDMM).
"""
import numpy as np
import torch

from . import api


class SparseSystem:
    """A CSR pattern in HBM whose values come from a tensor at each call."""

    def __init__(self, ctx, ptr, col, n=None, spd=False, solver="gmres", inner="NO", params=None,
                 adjoint_solver=None, adjoint_inner=None, adjoint_params=None, fp32=False):
        self.ctx = ctx
        ptr = np.ascontiguousarray(ptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        self.n = len(ptr) - 1 if n is None else n
        self.nnz = len(col)
        self.ptr, self.col = ptr, col
        self.A = api.Matrix(ctx, ptr, col, np.zeros(self.nnz), transpose=not spd, fp32=fp32, spd=spd)
        self.solver, self.inner = solver, inner
        self.params = dict(params or {})
        self.adjoint_solver = adjoint_solver or ("cg" if spd else solver)
        self.adjoint_inner = adjoint_inner if adjoint_inner is not None else ("NO" if spd else inner)
        self.adjoint_params = dict(adjoint_params if adjoint_params is not None else self.params)
        self._vals = None
        self.last_forward = None
        self.last_backward = None

    def _set(self, vals):
        # always refresh: the tensor may have been modified in place
        self.A.update_values(vals.detach().contiguous())
        self._vals = vals

    def solve(self, vals, b):
        return KrylovSolve.apply(vals, b, self)


class KrylovSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, b, system):
        if not (vals.is_cuda and b.is_cuda):
            raise TypeError("KrylovSolve expects device tensors")
        system._set(vals)
        x, res, _ = api.solve(system.ctx, system.A, b.detach().contiguous(), system.solver, system.inner,
                              history=False, **system.params)
        system.last_forward = res
        if res["status"] != 0:
            raise RuntimeError(f"forward solve did not converge: {res}")
        ctx.system = system
        ctx.save_for_backward(vals, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        vals, x = ctx.saved_tensors
        system = ctx.system
        system._set(vals)
        gb, gv, res = api.adjoint(system.ctx, system.A, x, gx.contiguous(), system.adjoint_solver,
                                  system.adjoint_inner, **system.adjoint_params)
        system.last_backward = res
        if res["status"] != 0:
            raise RuntimeError(f"adjoint solve did not converge: {res}")
        return gv, gb, None


class CentralFD:
    """Central difference (L(v + eps d) - L(v - eps d)) / (2 eps) of
    L = 0.5 ||x(v) - x_target||^2 along directions d, evaluated without
    cancellation.

    With A+ = A(v + eps d), A- = A(v - eps d) and D = A(d) (same pattern):
    x+ - x- = -2 eps A+^{-1} D x-  (exact identity), so with u = A+^{-1} D x-
    and x+ = x- - 2 eps u the difference quotient is
        -(u . (x- - x_target)) + eps (u . u),
    computed from two FP64 solves whose errors do not cancel catastrophically.
    The three matrices are built once; each direction refreshes their values
    on the device.
    """

    def __init__(self, ctx, ptr, col, vals, b, x_target, eps, solver, inner="NO", params=None, spd=False):
        self.ctx, self.b, self.xt, self.eps = ctx, b, x_target, eps
        self.solver, self.inner, self.params = solver, inner, dict(params or {})
        self.vals = torch.as_tensor(vals).detach().double().cuda().contiguous()
        fp32 = self.params.get("inner_precision", 64) == 32
        v = self.vals.cpu().numpy()
        self.Am = api.Matrix(ctx, ptr, col, v, transpose=not spd, spd=spd, fp32=fp32)
        self.Ap = api.Matrix(ctx, ptr, col, v, transpose=not spd, spd=spd, fp32=fp32)
        self.D = api.Matrix(ctx, ptr, col, v)

    def __call__(self, dvals, naive=False):
        d = torch.as_tensor(dvals).detach().double().cuda().contiguous()
        self.Am.update_values((self.vals - self.eps * d).contiguous())
        self.Ap.update_values((self.vals + self.eps * d).contiguous())
        self.D.update_values(d)
        kw = dict(history=False, **self.params)
        xm, rm, _ = api.solve(self.ctx, self.Am, self.b, self.solver, self.inner, **kw)
        u, ru, _ = api.solve(self.ctx, self.Ap, self.D.spmv(xm), self.solver, self.inner, **kw)
        for r in (rm, ru):
            if r["status"] != 0:
                raise RuntimeError(f"FD solve did not converge: {r}")
        fd = float(-(u @ (xm - self.xt)) + self.eps * (u @ u))
        if not naive:
            return fd
        xp, rp, _ = api.solve(self.ctx, self.Ap, self.b, self.solver, self.inner, **kw)
        nv = float((0.5 * ((xp - self.xt) ** 2).sum() - 0.5 * ((xm - self.xt) ** 2).sum()) / (2 * self.eps))
        return fd, nv

    def close(self):
        for m in (self.Am, self.Ap, self.D):
            m.close()


def central_fd(ctx, ptr, col, vals, dvals, b, x_target, eps, solver, inner="NO", params=None, spd=False,
               naive=True):
    """One CentralFD evaluation: (fd, naive_fd) where naive_fd subtracts the
    two losses directly (a third solve); with naive=False returns fd only."""
    f = CentralFD(ctx, ptr, col, vals, b, x_target, eps, solver, inner, params, spd)
    try:
        return f(dvals, naive=naive)
    finally:
        f.close()
