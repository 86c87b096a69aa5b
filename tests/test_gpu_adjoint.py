"""Adjoint (torch.autograd) checks on the GPU: against the FP64 dense
reference gradient and against central finite differences (FP64).

Tolerance: the adjoint gradient must agree with central FD to <= 1e-6
relative in every direction.  Central FD carries a truncation error of order
eps^2 * ||A^-1 D||^2 (D the value perturbation); for the Zipf matrix below
(values spanning 1.5e-5..34) that is ~1e-6 at eps=1e-6 and falls 100x per 10x
smaller eps (checked with dense FP64 solves).  The FD here takes x+ - x- from
the resolvent identity, which has no cancellation, so eps=1e-7 is used: the
remaining FD error is ~1e-8 relative, set by the 1e-14 solves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dense_grad(ptr, col, val, x, g):
    from tests._problems import dense
    D = dense(ptr, col, val)
    lam = np.linalg.solve(D.T, g)
    rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    return lam, -lam[rows] * x[col]


def test_spd_vpcg_forward_cg_adjoint(ctx):
    import torch
    from solver_cpp_amd import autograd as AG
    from solver_cpp_amd import gen
    g = 32
    ptr, col, val, kap = gen.aniso2d(g, g, 4)
    n = g * g
    b = torch.from_numpy(gen.vector(41, n, "uniform")).cuda()
    xt = torch.from_numpy(gen.vector(40, n, "normal")).cuda()
    sysm = AG.SparseSystem(ctx, ptr, col, spd=True, solver="vpcg", inner="cg", fp32=True,
                           params=dict(outer_eps=1e-14, outer_maxloop=3000, inner_maxloop=10, inner_precision=32),
                           adjoint_solver="cg", adjoint_params=dict(outer_eps=1e-14, outer_maxloop=5000))
    vals = torch.from_numpy(val).cuda().requires_grad_(True)
    bb = b.clone().requires_grad_(True)
    x = sysm.solve(vals, bb)
    L = 0.5 * ((x - xt) ** 2).sum()
    L.backward()
    assert sysm.last_forward["status"] == 0 and sysm.last_backward["status"] == 0
    lam, gv = _dense_grad(ptr, col, val, x.detach().cpu().numpy(), (x - xt).detach().cpu().numpy())
    assert np.linalg.norm(bb.grad.cpu().numpy() - lam) <= 1e-9 * np.linalg.norm(lam)
    assert np.linalg.norm(vals.grad.cpu().numpy() - gv) <= 1e-9 * np.linalg.norm(gv)
    # FD along symmetric edge directions (keeps A SPD): d_val = A(d_kappa)
    rng = np.random.default_rng(42)
    for _ in range(6):
        dk = rng.standard_normal(len(kap))
        dv = torch.from_numpy(gen.aniso2d_values(g, g, dk)).cuda()
        fd, naive = AG.central_fd(ctx, ptr, col, vals.detach(), dv, b, xt, 1e-6, "cg",
                                  params=dict(outer_eps=1e-14, outer_maxloop=5000), spd=True)
        ad = float(vals.grad @ dv)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad, naive)


@pytest.mark.parametrize("case", ["convdiff", "zipf"])
def test_nonsymmetric_vpgmres_forward_gmres_adjoint(ctx, case):
    import torch
    from solver_cpp_amd import autograd as AG
    from solver_cpp_amd import gen
    if case == "convdiff":
        ptr, col, val = gen.stencil(3, 14, 14, 14, pe=0.8)
    else:
        # cap 32: the full Zipf cap (4096) spreads row scales ~3600x and plain
        # GMRES needs >10^4 steps to 1e-14 even in FP64 (oracle)
        ptr, col, val = gen.zipf(3000, 1.5, 32, (51, 52, 53))
    n = len(ptr) - 1
    b = torch.from_numpy(gen.vector(50, n, "uniform")).cuda()
    xt = torch.from_numpy(gen.vector(5, n, "normal")).cuda()
    fwd = dict(outer_eps=1e-14, outer_maxloop=3000, outer_restart=21, inner_maxloop=11, inner_restart=12,
               inner_precision=32)
    sysm = AG.SparseSystem(ctx, ptr, col, spd=False, solver="vpgmres", inner="gmres", fp32=True, params=fwd,
                           adjoint_solver="gmres", adjoint_inner="NO",
                           adjoint_params=dict(outer_eps=1e-14, outer_maxloop=20000, outer_restart=41))
    vals = torch.from_numpy(val).cuda().requires_grad_(True)
    bb = b.clone().requires_grad_(True)
    x = sysm.solve(vals, bb)
    L = 0.5 * ((x - xt) ** 2).sum()
    L.backward()
    lam, gv = _dense_grad(ptr, col, val, x.detach().cpu().numpy(), (x - xt).detach().cpu().numpy())
    assert np.linalg.norm(bb.grad.cpu().numpy() - lam) <= 1e-9 * np.linalg.norm(lam)
    assert np.linalg.norm(vals.grad.cpu().numpy() - gv) <= 1e-9 * np.linalg.norm(gv)
    rng = np.random.default_rng(54)
    for _ in range(6):
        dv = torch.from_numpy(rng.standard_normal(len(val))).cuda()
        fd, naive = AG.central_fd(ctx, ptr, col, vals.detach(), dv, b, xt, 1e-7, "vpgmres", "gmres",
                                  params=fwd)
        ad = float(vals.grad @ dv)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad, naive)


def test_zipf_heavy_rows_row_scaled(ctx):
    # Zipf rows up to 2048 entries: diagonals spread over decades and
    # restarted GMRES on A stagnates; on D^-1 A (SparseSystem(row_scale=True))
    # mixed VPGMRES and the adjoint GMRES converge, and the gradient is that of
    # the unscaled system: the dense FP64 gradient and central FD
    import torch
    from solver_cpp_amd import autograd as AG
    from solver_cpp_amd import gen
    ptr, col, val = gen.zipf(3000, 1.5, 2048, (51, 52, 53))
    n = len(ptr) - 1
    assert np.diff(ptr).max() > 1000
    b = torch.from_numpy(gen.vector(50, n, "uniform")).cuda()
    xt = torch.from_numpy(gen.vector(5, n, "normal")).cuda()
    fwd = dict(outer_eps=1e-14, outer_maxloop=500, outer_restart=21, inner_maxloop=11, inner_restart=12,
               inner_precision=32)
    sysm = AG.SparseSystem(ctx, ptr, col, spd=False, solver="vpgmres", inner="gmres", fp32=True, params=fwd,
                           adjoint_solver="gmres", adjoint_params=dict(outer_eps=1e-14, outer_maxloop=2000,
                                                                       outer_restart=21), row_scale=True)
    vals = torch.from_numpy(val).cuda().requires_grad_(True)
    bb = b.clone().requires_grad_(True)
    x = sysm.solve(vals, bb)
    L = 0.5 * ((x - xt) ** 2).sum()
    L.backward()
    assert sysm.last_forward["status"] == 0 and sysm.last_backward["status"] == 0
    assert sysm.last_forward["iters"] < 50 and sysm.last_backward["iters"] < 200
    lam, gv = _dense_grad(ptr, col, val, x.detach().cpu().numpy(), (x - xt).detach().cpu().numpy())
    assert np.linalg.norm(bb.grad.cpu().numpy() - lam) <= 1e-9 * np.linalg.norm(lam)
    assert np.linalg.norm(vals.grad.cpu().numpy() - gv) <= 1e-9 * np.linalg.norm(gv)
    rng = np.random.default_rng(54)
    fdx = AG.CentralFD(ctx, ptr, col, vals.detach(), b, xt, 1e-7, "vpgmres", "gmres", params=fwd, row_scale=True)
    for _ in range(6):
        dv = torch.from_numpy(rng.standard_normal(len(val))).cuda()
        fd, naive = fdx(dv, naive=True)
        ad = float(vals.grad @ dv)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad, naive)
    fdx.close()

