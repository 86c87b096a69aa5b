"""The benchmark workloads at their full sizes, checked through properties
that hold at any size (the oracle's own runs are kept to small sizes).

- 3-D 7-point Poisson 160^3 (n = 4,096,000): mixed VPGMRES(20) with ten FP32
  GMRES inner steps to 1e-10 (vpgmres.cpp:746-975 with the inner solve of
  innerMethods.cpp:71-118).  Checked: the true residual, recomputed on the
  host with scipy's CSR product, meets the tolerance; the solve is
  bit-identical run to run; and x lies within kappa(A) * 1e-10 of the FP64
  GMRES solution, kappa(A) = lambda_max / lambda_min of the 7-point stencil
  in closed form.
- SpMV (cudaFunction.cu:127-156, blas.cpp:43-129) and the explicit A^T
  (cudaFunction.cu:896-931) at the full sizes of the 160^3 Poisson, 192^3
  convection-diffusion and 4,000,000-row Zipf matrices (the last on column
  panels): FP64 and FP32 against scipy's FP64 product within the row-sum
  rounding bound, the adjoint identity (z, A x) = (A^T z, x), and A^T equal
  entry for entry to scipy's stable CSC -> CSR conversion.
"""
import math

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
