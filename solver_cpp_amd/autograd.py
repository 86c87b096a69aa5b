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

Row scaling (SparseSystem(row_scale=True)): the solves run on the
equilibrated system D^-1 A x = D^-1 b, D = diag(A), whose rows all have unit
diagonal (for a matrix whose row scales spread over decades, where restarted
Krylov methods on A itself stagnate).  x is the same; the adjoint solves
(D^-1 A)^T mu = dL/dx on the explicit transpose and maps back: lambda =
D^-1 mu, dL/dA_ij = -mu_i x_j / d_i.  D only changes how the systems are
solved, not the function x(A, b), so the gradient is that of the unscaled
system.
"""
import numpy as np
import torch

from . import api


class RowScale:
    """Row index of every entry and the slot of every row's diagonal entry of
    a CSR pattern (on the device), for D = diag(A) row scaling."""

    def __init__(self, ptr, col, device="cuda"):
        n = len(ptr) - 1
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ptr))
        on_diag = np.flatnonzero(col == rows)
        dslot = np.full(n, -1, np.int64)
        dslot[rows[on_diag]] = on_diag
        if (dslot < 0).any():
            raise ValueError(f"row scaling needs a diagonal entry in every row "
                             f"(row {int(np.flatnonzero(dslot < 0)[0])} has none)")
        self.rows = torch.from_numpy(rows).to(device)
        self.dslot = torch.from_numpy(dslot).to(device)

    def diag(self, vals):
        d = vals.detach()[self.dslot]
        if not bool(torch.all(torch.isfinite(d) & (d != 0))):
            raise ValueError("row scaling needs a finite nonzero diagonal in every row")
        return d

    def scaled(self, vals, d):
        """values of D^-1 A"""
        return (vals.detach() / d[self.rows]).contiguous()


class SparseSystem:
    """A CSR pattern in HBM whose values come from a tensor at each call.

    row_scale=True solves the row-equilibrated system D^-1 A x = D^-1 b
    (D = diag(A); every row needs a nonzero diagonal entry), see the module
    docstring."""

    def __init__(self, ctx, ptr, col, n=None, spd=False, solver="gmres", inner="NO", params=None,
                 adjoint_solver=None, adjoint_inner=None, adjoint_params=None, fp32=False, row_scale=False):
        self.ctx = ctx
        ptr = np.ascontiguousarray(ptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        self.n = len(ptr) - 1 if n is None else n
        self.nnz = len(col)
        self.ptr, self.col = ptr, col
        if row_scale and spd:
            raise ValueError("row scaling makes the matrix nonsymmetric: use it with spd=False")
        self.scale = RowScale(ptr, col) if row_scale else None
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
        """Load the values (of D^-1 A with row scaling); returns D's diagonal
        (None without row scaling).  Always refreshes: the tensor may have
        been modified in place."""
        self._vals = vals
        if self.scale is None:
            self.A.update_values(vals.detach().contiguous())
            return None
        d = self.scale.diag(vals)
        self.A.update_values(self.scale.scaled(vals, d))
        return d

    def solve(self, vals, b):
        return KrylovSolve.apply(vals, b, self)


class KrylovSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, b, system):
        if not (vals.is_cuda and b.is_cuda):
            raise TypeError("KrylovSolve expects device tensors")
        d = system._set(vals)
        rhs = b.detach().contiguous() if d is None else (b.detach() / d).contiguous()
        x, res, _ = api.solve(system.ctx, system.A, rhs, system.solver, system.inner, history=False, **system.params)
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
        d = system._set(vals)
        gb, gv, res = api.adjoint(system.ctx, system.A, x, gx.contiguous(), system.adjoint_solver,
                                  system.adjoint_inner, **system.adjoint_params)
        system.last_backward = res
        if res["status"] != 0:
            raise RuntimeError(f"adjoint solve did not converge: {res}")
        if d is not None:
            # (D^-1 A)^T mu = dL/dx gave mu and -mu_i x_j: lambda = D^-1 mu
            gb = gb / d
            gv = gv / d[system.scale.rows]
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

    def __init__(self, ctx, ptr, col, vals, b, x_target, eps, solver, inner="NO", params=None, spd=False,
                 row_scale=False):
        self.ctx, self.b, self.xt, self.eps = ctx, b, x_target, eps
        self.solver, self.inner, self.params = solver, inner, dict(params or {})
        self.vals = torch.as_tensor(vals).detach().double().cuda().contiguous()
        # row_scale: each solve runs on its matrix's D^-1 A (see SparseSystem)
        self.scale = RowScale(ptr, col) if row_scale else None
        fp32 = self.params.get("inner_precision", 64) == 32
        v = self.vals.cpu().numpy()
        self.Am = api.Matrix(ctx, ptr, col, v, transpose=not spd, spd=spd, fp32=fp32)
        self.Ap = api.Matrix(ctx, ptr, col, v, transpose=not spd, spd=spd, fp32=fp32)
        self.D = api.Matrix(ctx, ptr, col, v)

    def __call__(self, dvals, naive=False):
        d = torch.as_tensor(dvals).detach().double().cuda().contiguous()
        vm = (self.vals - self.eps * d).contiguous()
        vp = (self.vals + self.eps * d).contiguous()
        self.D.update_values(d)
        kw = dict(history=False, **self.params)
        if self.scale is None:
            self.Am.update_values(vm)
            self.Ap.update_values(vp)
            xm, rm, _ = api.solve(self.ctx, self.Am, self.b, self.solver, self.inner, **kw)
            u, ru, _ = api.solve(self.ctx, self.Ap, self.D.spmv(xm), self.solver, self.inner, **kw)
        else:
            dm, dp = self.scale.diag(vm), self.scale.diag(vp)
            self.Am.update_values(self.scale.scaled(vm, dm))
            self.Ap.update_values(self.scale.scaled(vp, dp))
            xm, rm, _ = api.solve(self.ctx, self.Am, (self.b / dm).contiguous(), self.solver, self.inner, **kw)
            u, ru, _ = api.solve(self.ctx, self.Ap, (self.D.spmv(xm) / dp).contiguous(), self.solver, self.inner,
                                 **kw)
        for r in (rm, ru):
            if r["status"] != 0:
                raise RuntimeError(f"FD solve did not converge: {r}")
        fd = float(-(u @ (xm - self.xt)) + self.eps * (u @ u))
        if not naive:
            return fd
        bp = self.b if self.scale is None else (self.b / dp).contiguous()
        xp, rp, _ = api.solve(self.ctx, self.Ap, bp, self.solver, self.inner, **kw)
        nv = float((0.5 * ((xp - self.xt) ** 2).sum() - 0.5 * ((xm - self.xt) ** 2).sum()) / (2 * self.eps))
        return fd, nv

    def close(self):
        for m in (self.Am, self.Ap, self.D):
            m.close()


def central_fd(ctx, ptr, col, vals, dvals, b, x_target, eps, solver, inner="NO", params=None, spd=False,
               naive=True, row_scale=False):
    """One CentralFD evaluation: (fd, naive_fd) where naive_fd subtracts the
    two losses directly (a third solve); with naive=False returns fd only."""
    f = CentralFD(ctx, ptr, col, vals, b, x_target, eps, solver, inner, params, spd, row_scale)
    try:
        return f(dvals, naive=naive)
    finally:
        f.close()
