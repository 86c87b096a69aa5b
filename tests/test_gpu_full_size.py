"""The benchmark workloads at their full sizes, checked through properties
that hold at any size (the oracle's own runs are kept to small sizes).

- 3-D 7-point Poisson 160^3 (n = 4,096,000): mixed VPGMRES(20) with ten FP32
  GMRES inner steps to 1e-10 (vpgmres.cpp:746-975 with the inner solve of
  innerMethods.cpp:71-118).  Checked: the true residual, recomputed on the
  host with scipy's CSR product, meets the tolerance; the solve is
  bit-identical run to run; and x lies within kappa(A) * 1e-10 of the FP64
  GMRES solution, kappa(A) = lambda_max / lambda_min of the 7-point stencil
  in closed form.
- The adjoint of the 4,000,000-row Zipf(1.5) matrix (rows capped at 4096
  entries, 196 M nonzeros): forward mixed VPGMRES(20) with ten FP32 GMRES
  inner steps, backward GMRES(20) on the explicit transpose, both on the
  row-equilibrated system D^-1 A (D = diag(A)), to 1e-14.  Checked: the
  unscaled true residuals of both solves, 64 central finite differences
  along relative value directions, and the oracle's FP64 adjoint solve of
  the same system on the host.
- The adjoints of the 1024^2 anisotropic diffusion matrix (SPD: forward
  VPCG with an FP32 inner CG, backward CG + SDDMM) and of the 192^3
  convection-diffusion matrix at Peclet 0.8 (forward mixed VPGMRES(20),
  backward GMRES(20) on the explicit transpose), to 1e-14: 64 central
  finite differences each, eps = 1e-6 (symmetric edge perturbations of the
  conductivities for the SPD case, which keep A SPD; seeded N(0,1) value
  directions for the other), every direction within 1e-6 relative.
- SpMV (cudaFunction.cu:127-156, blas.cpp:43-129) and the explicit A^T
  (cudaFunction.cu:896-931) at the full sizes of the 160^3 Poisson, 192^3
  convection-diffusion and 4,000,000-row Zipf matrices (the last on column
  panels): FP64 and FP32 against scipy's FP64 product within the row-sum
  rounding bound, the adjoint identity (z, A x) = (A^T z, x), and A^T equal
  entry for entry to scipy's stable CSC -> CSR conversion.
"""
import math
import time

import numpy as np
import pytest

from tests import _problems as P

pytestmark = pytest.mark.gpu

EPS64 = 2.0 ** -53
EPS32 = 2.0 ** -24


def _scipy(ptr, col, val):
    import scipy.sparse as sp
    n = len(ptr) - 1
    return sp.csr_matrix((val, col, ptr), shape=(n, n))


def test_headline_solve_full_size(ctx):
    import torch
    from solver_cpp_amd import api, gen
    g = 160
    Pb = gen.problem("poisson3d")
    ptr, col, val = Pb["ptr"], Pb["col"], Pb["val"]
    assert len(ptr) - 1 == g ** 3
    A = api.Matrix(ctx, ptr, col, val, fp32=True)
    b = torch.from_numpy(Pb["b"]).cuda()
    mixed = dict(outer_restart=21, outer_maxloop=5000, outer_eps=1e-10, inner_maxloop=11, inner_restart=12,
                 inner_eps=0.0, inner_precision=32)
    x1, r1, _ = api.solve(ctx, A, b, "vpgmres", "gmres", history=False, **mixed)
    x2, r2, _ = api.solve(ctx, A, b, "vpgmres", "gmres", history=False, **mixed)
    assert r1["status"] == 0 and r1["iters"] == r2["iters"]
    assert torch.equal(x1, x2)  # fixed-order reductions: run to run bit-identical
    # true residual on the host, independent of the library's SpMV
    S = _scipy(ptr, col, val)
    xh, bh = x1.cpu().numpy(), Pb["b"]
    rel = np.linalg.norm(bh - S @ xh) / np.linalg.norm(bh)
    assert rel <= 1e-10, rel
    # the library's own FP64 residual: the same up to the rounding of b - A x
    assert abs(rel - r1["relres_true"]) <= 1e-2 * rel, (rel, r1["relres_true"])
    # FP64 GMRES(20) to 1e-12; |x - x*| / |x*| <= kappa * relres for both
    x64, r64, _ = api.solve(ctx, A, b, "gmres", history=False, outer_restart=21, outer_maxloop=400000,
                            outer_eps=1e-12)
    assert r64["status"] == 0
    lam_min = 6.0 - 6.0 * math.cos(math.pi / (g + 1))
    lam_max = 6.0 + 6.0 * math.cos(math.pi / (g + 1))
    kappa = lam_max / lam_min
    err = float(torch.linalg.norm(x1 - x64) / torch.linalg.norm(x64))
    assert err <= kappa * (1e-10 + 1e-12), (err, kappa)
    A.close()


@pytest.mark.parametrize("cfg", ["poisson3d", "convdiff3d_pe08", "zipf"])
def test_spmv_full_size(ctx, cfg):
    from solver_cpp_amd import api, gen
    Pb = gen.problem(cfg)
    ptr, col, val = Pb["ptr"], Pb["col"], Pb["val"]
    n = len(ptr) - 1
    A = api.Matrix(ctx, ptr, col, val, transpose=True, fp32=True)
    if cfg == "zipf":
        assert A.info()["spmv_kind"] == 3  # column panels, chosen by the library
    m = int(np.diff(ptr).max())
    S = _scipy(ptr, col, val)
    x = np.random.default_rng(61).standard_normal(n)
    z = np.random.default_rng(62).standard_normal(n)
    bnd = P.row_abs_bound(ptr, col, val, x)
    y = A.spmv(x)
    assert np.all(np.abs(y - S @ x) <= 4 * m * EPS64 * bnd + 1e-300)
    # A^T against scipy's transposed product, and the adjoint identity
    mt = int(np.bincount(col, minlength=n).max())
    yt = A.spmv(z, trans=True)
    # scipy's CSC -> CSR conversion is a stable counting sort: the A^T the
    # device built must equal it entry for entry (repeated entries kept)
    St = S.T.tocsr()
    tp, tc, tv = A.transpose_csr()
    assert np.array_equal(tp, St.indptr) and np.array_equal(tc, St.indices) and np.array_equal(tv, St.data)
    assert np.all(np.abs(yt - St @ z) <= 4 * mt * EPS64 * (abs(St) @ np.abs(z)) + 1e-300)
    # exact sums of the rounded products: what is left is the two SpMVs'
    # rounding, (m + 1) and (mt + 1) eps times sum |z_i a_ij x_j|, and the
    # two final roundings
    lhs, rhs = math.fsum(z * y), math.fsum(yt * x)
    assert abs(lhs - rhs) <= 2 * (m + mt + 4) * EPS64 * (np.abs(z) @ bnd), (lhs, rhs)
    # FP32 values and vectors against the FP64 product of the same FP32 inputs
    x32, v32 = x.astype(np.float32), val.astype(np.float32)
    y32 = A.spmv(x32, precision=32)
    S32 = _scipy(ptr, col, v32.astype(np.float64))
    b32 = 4 * m * EPS32 * P.row_abs_bound(ptr, col, v32.astype(np.float64), x32.astype(np.float64)) + 1e-30
    assert np.all(np.abs(y32.astype(np.float64) - S32 @ x32.astype(np.float64)) <= b32)
    A.close()


def test_zipf_adjoint_full_size(ctx, oracle):
    import torch
    from solver_cpp_amd import api, gen
    from solver_cpp_amd import autograd as AG
    t0 = time.perf_counter()
    Pb = gen.problem("zipf")
    ptr, col, val = Pb["ptr"], Pb["col"], Pb["val"]
    n, nnz = len(ptr) - 1, len(col)
    assert n == 4_000_000 and np.diff(ptr).max() == 4096
    print(f"[zipf-adjoint] generated n={n} nnz={nnz} in {time.perf_counter() - t0:.1f} s", flush=True)
    b = torch.from_numpy(Pb["b"]).cuda()
    xt = torch.from_numpy(Pb["x_target"]).cuda()
    fwd = dict(outer_eps=1e-14, outer_maxloop=400, outer_restart=21, inner_maxloop=11, inner_restart=12,
               inner_eps=0.0, inner_precision=32)
    adj = dict(outer_eps=1e-14, outer_maxloop=2000, outer_restart=21)
    sysm = AG.SparseSystem(ctx, ptr, col, spd=False, solver="vpgmres", inner="gmres", fp32=True, params=fwd,
                           adjoint_solver="gmres", adjoint_params=adj, row_scale=True)
    vals = torch.from_numpy(val).cuda().requires_grad_(True)
    bb = b.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    x = sysm.solve(vals, bb)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    L = 0.5 * ((x - xt) ** 2).sum()
    L.backward()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    rf, rb = sysm.last_forward, sysm.last_backward
    print(f"[zipf-adjoint] forward {t2 - t1:.3f} s ({rf['iters']} it), backward {t3 - t2:.3f} s "
          f"({rb['iters']} it)", flush=True)
    assert rf["status"] == 0 and rb["status"] == 0
    # unscaled true residuals with the library's FP64 SpMV of A and A^T
    A = api.Matrix(ctx, ptr, col, val, transpose=True)
    g = (x - xt).detach()
    lam = bb.grad
    fr = float(torch.linalg.norm(b - A.spmv(x.detach())) / torch.linalg.norm(b))
    br = float(torch.linalg.norm(A.spmv(lam, trans=True) - g) / torch.linalg.norm(g))
    print(f"[zipf-adjoint] unscaled relres forward {fr:.2e}, adjoint {br:.2e}", flush=True)
    assert fr <= 1e-11 and br <= 1e-11
    A.close()
    # the oracle's FP64 GMRES(20) on (D^-1 A)^T from the same x and g
    rows = np.repeat(np.arange(n), np.diff(ptr))
    dg = np.zeros(n)
    on = col == rows
    dg[rows[on]] = val[on]
    gh = g.cpu().numpy()
    mu, _, ro = oracle.adjoint(ptr, col, val / dg[rows], x.detach().cpu().numpy(), gh, "gmres",
                               omp_threads=min(oracle.max_threads(), 16), **adj)
    lo = mu / dg
    assert ro["status"] == 0
    lg = lam.cpu().numpy()
    rel = np.linalg.norm(lg - lo) / np.linalg.norm(lo)
    print(f"[zipf-adjoint] grad_b vs oracle {rel:.2e} ({ro['iters']} oracle it)", flush=True)
    assert rel <= 1e-9
    # 64 central finite differences (eps 1e-7) along seeded relative value
    # directions d_k = a_k z_k, z ~ N(0,1): every entry is perturbed in
    # proportion to its own size.  (Absolute N(0,1) perturbations of 1e-7 are
    # comparable to the smallest diagonals here, where a row's only
    # off-diagonal entry is tiny, and the difference quotient's O(eps^2)
    # term then swamps it.)
    gen_d = torch.Generator(device="cuda")
    gen_d.manual_seed(54)
    v0 = vals.detach()
    fdx = AG.CentralFD(ctx, ptr, col, v0, b, xt, 1e-7, "vpgmres", "gmres", params=fwd, row_scale=True)
    worst = 0.0
    t4 = time.perf_counter()
    for i in range(64):
        dv = v0 * torch.randn(nnz, dtype=torch.float64, device="cuda", generator=gen_d)
        fd = fdx(dv)
        ad = float(vals.grad @ dv)
        worst = max(worst, abs(fd - ad) / abs(ad))
    fdx.close()
    print(f"[zipf-adjoint] 64 FD directions in {time.perf_counter() - t4:.1f} s, max rel err {worst:.2e}",
          flush=True)
    assert worst <= 1e-6
    sysm.A.close()


@pytest.mark.parametrize("cfg", ["aniso2d", "convdiff3d_pe08"])
def test_adjoint_fd_full_size(ctx, cfg):
    import torch
    from solver_cpp_amd import autograd as AG
    from solver_cpp_amd import gen
    Pb = gen.problem(cfg)
    ptr, col, val = Pb["ptr"], Pb["col"], Pb["val"]
    nnz = len(col)
    b = torch.from_numpy(Pb["b"]).cuda()
    xt = torch.from_numpy(Pb["x_target"]).cuda()
    if cfg == "aniso2d":
        fwd = dict(outer_eps=1e-14, outer_maxloop=20000, inner_maxloop=10, inner_precision=32)
        cg = dict(outer_eps=1e-14, outer_maxloop=100000)
        sysm = AG.SparseSystem(ctx, ptr, col, spd=True, solver="vpcg", inner="cg", fp32=True, params=fwd,
                               adjoint_solver="cg", adjoint_params=dict(cg))
        fd_args = ("cg", "NO", cg)
    else:
        fwd = dict(outer_eps=1e-14, outer_maxloop=20000, outer_restart=21, inner_maxloop=11, inner_restart=12,
                   inner_eps=0.0, inner_precision=32)
        sysm = AG.SparseSystem(ctx, ptr, col, spd=False, solver="vpgmres", inner="gmres", fp32=True, params=fwd,
                               adjoint_solver="gmres",
                               adjoint_params=dict(outer_eps=1e-14, outer_maxloop=100000, outer_restart=21))
        fd_args = ("vpgmres", "gmres", fwd)
    vals = torch.from_numpy(val).cuda().requires_grad_(True)
    bb = b.clone().requires_grad_(True)
    x = sysm.solve(vals, bb)
    (0.5 * ((x - xt) ** 2).sum()).backward()
    assert sysm.last_forward["status"] == 0 and sysm.last_backward["status"] == 0
    spd = cfg == "aniso2d"
    fdx = AG.CentralFD(ctx, ptr, col, vals.detach(), b, xt, 1e-6, fd_args[0], fd_args[1], params=fd_args[2],
                       spd=spd)
    rng = np.random.default_rng(42 if spd else 54)
    worst = 0.0
    t0 = time.perf_counter()
    for _ in range(64):
        if spd:
            dv = torch.from_numpy(gen.aniso2d_values(Pb["grid"], Pb["grid"], rng.standard_normal(len(Pb["kappa"]))))
        else:
            dv = torch.from_numpy(rng.standard_normal(nnz))
        dv = dv.cuda()
        fd = fdx(dv)
        ad = float(vals.grad @ dv)
        worst = max(worst, abs(fd - ad) / abs(ad))
    fdx.close()
    print(f"[adjoint-fd] {cfg}: forward {sysm.last_forward['iters']} it, adjoint {sysm.last_backward['iters']} it, "
          f"64 FD directions in {time.perf_counter() - t0:.1f} s, max rel err {worst:.2e}", flush=True)
    assert worst <= 1e-6
    sysm.A.close()

