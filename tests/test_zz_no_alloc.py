"""ks_solve_noalloc (plain FP64 CG with its state in the library's fixed
device scratch array) against ks_solve on the same inputs, through the C ABI:
two calls in a row, the second on a smaller matrix, each checked against
ks_solve's CG (same status and iteration count, x equal to within 1e-12
relative: the two run the same kernels in the same order), plus the size
limit and the pointer-location rule."""
import ctypes as C

import numpy as np
import pytest

from solver_cpp_amd import _abi
from solver_cpp_amd._abi import KsResult, make_params

pytestmark = pytest.mark.gpu


def _noalloc(ctx, A, b, **kw):
    import torch
    p = make_params("cg", "NO", **kw)
    p.flags |= _abi.KS_PTR_DEVICE
    x = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    res = KsResult()
    torch.cuda.synchronize()
    st = ctx.lib.ks_solve_noalloc(ctx.h, A.h, C.byref(p), b.data_ptr(), x.data_ptr(), C.byref(res))
    return st, x, res.as_dict()


@pytest.mark.parametrize("first,second", [(("poisson2d", 256), ("poisson3d_sym", 24)),
                                          (("aniso2d", 1024), ("poisson2d", 48))])
def test_noalloc_cg_matches_ks_solve(ctx, first, second):
    import torch
    from solver_cpp_amd import api, gen
    kw = dict(outer_eps=1e-12, outer_maxloop=30000)
    for cfg, g in (first, second):  # the second call is on a smaller matrix
        print(f"[noalloc] {cfg} {g}", flush=True)
        P = gen.problem(cfg, scale=g)
        A = api.Matrix(ctx, P["ptr"], P["col"], P["val"])
        b = torch.from_numpy(P["b"]).cuda()
        xr, rr, _ = api.solve(ctx, A, b, "cg", history=False, **kw)
        st, x, r = _noalloc(ctx, A, b, **kw)
        assert st == 0, ctx.lib.ks_ctx_last_error(ctx.h)
        assert r["status"] == rr["status"] == 0 and r["iters"] == rr["iters"], (r["iters"], rr["iters"])
        assert float(torch.linalg.norm(x - xr) / torch.linalg.norm(xr)) <= 1e-12
        assert abs(r["relres_true"] - rr["relres_true"]) <= 1e-3 * rr["relres_true"]
        assert r["spmv"] == rr["spmv"] and r["steps"] == rr["steps"]
        A.close()


def test_noalloc_limits(ctx):
    import torch
    from solver_cpp_amd import api, gen
    # one row past the maximum: rejected before any kernel runs
    ptr, col, val = gen.stencil(2, 1025, 1024)
    assert len(ptr) - 1 == (1 << 20) + 1024
    A = api.Matrix(ctx, ptr, col, val)
    b = torch.ones(A.n, dtype=torch.float64, device="cuda")
    st, x, r = _noalloc(ctx, A, b, outer_eps=1e-8)
    assert st == _abi.KS_ERR_BAD_INPUT
    assert "exceeds its maximum" in ctx.lib.ks_ctx_last_error(ctx.h).decode()
    A.close()
    # host arrays and other solvers are refused
    ptr, col, val = gen.stencil(2, 8, 8)
    A = api.Matrix(ctx, ptr, col, val)
    p = make_params("cg", "NO")
    res = KsResult()
    bh, xh = np.ones(64), np.zeros(64)
    assert ctx.lib.ks_solve_noalloc(ctx.h, A.h, C.byref(p), bh.ctypes.data, xh.ctypes.data,
                                    C.byref(res)) == _abi.KS_ERR_BAD_INPUT
    p = make_params("gmres", "NO")
    p.flags |= _abi.KS_PTR_DEVICE
    bd = torch.ones(64, dtype=torch.float64, device="cuda")
    xd = torch.zeros(64, dtype=torch.float64, device="cuda")
    assert ctx.lib.ks_solve_noalloc(ctx.h, A.h, C.byref(p), bd.data_ptr(), xd.data_ptr(),
                                    C.byref(res)) == _abi.KS_ERR_UNSUPPORTED
    A.close()
