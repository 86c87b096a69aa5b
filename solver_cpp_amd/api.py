"""Python mirror of the reference's Krylov operator / solver interface over the
libksolve C ABI.

Reference interface mirrored (file:line in ikuno/solver_cpp):
  * solver names and flags   -> collection::readCMD (solver_collection.cpp:269-387)
  * outer / inner dispatch   -> outerMethods::outerSelect (outerMethods.cpp:20-82),
                                innerMethods::innerSelect (innerMethods.cpp:71-118)
  * device operators         -> class cuda (cudaFunction.hpp:7-129): MtxVec_mult,
                                CSR2CSC, Kskip_cg_bicg_base
  * host BLAS                -> class blas (blas.hpp:8-78): dot, norm_2,
                                Scalar_axy, Gmres_sp_1, Kskip_cg_innerProduce
Return values follow the reference: solve() returns the solver's exit flag
(0 converge, 2 not converge) in result["status"] and the `loop =` count in
result["iters"]; errors raise KsolveError instead of exiting the process.

Arrays may be numpy arrays (host memory; copied to and from HBM) or torch
tensors on the GPU (used in place).
"""
import ctypes as C
import weakref

import numpy as np

from . import _abi
from ._abi import KsParams, KsResult, make_params
from ._lib import load

try:  # torch is optional for host-array use
    import torch
except Exception:  # pragma: no cover
    torch = None


class KsolveError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[{status}] {msg}")
        self.status = status


def _is_dev(a):
    return torch is not None and isinstance(a, torch.Tensor) and a.is_cuda


def _torch_dtype(dt):
    return {np.float64: torch.float64, np.float32: torch.float32, np.int32: torch.int32}[dt]


class _Arg:
    """Pointer for one array argument; keeps the buffer alive."""

    def __init__(self, a, dtype, writable=False):
        self.dev = _is_dev(a)
        if self.dev:
            if a.dtype != _torch_dtype(dtype) or not a.is_contiguous():
                raise TypeError(f"device tensor must be contiguous {dtype.__name__}")
            self.obj = a
            self.ptr = a.data_ptr()
        else:
            arr = np.asarray(a)
            if writable:
                if arr.dtype != dtype or not arr.flags["C_CONTIGUOUS"] or not arr.flags["WRITEABLE"]:
                    raise TypeError(f"output array must be a writable contiguous {dtype.__name__} array")
            else:
                arr = np.ascontiguousarray(arr, dtype=dtype)
            self.obj = arr
            self.ptr = arr.ctypes.data


def _flags(*args):
    devs = {a.dev for a in args if a is not None}
    if len(devs) > 1:
        raise TypeError("mix of host arrays and device tensors in one call")
    if devs == {True}:
        torch.cuda.synchronize()
        return _abi.KS_PTR_DEVICE
    return 0


class Context:
    """One GPU, one HIP stream (cuda::cuda + collection set-up)."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.ks_ctx_create(device, C.byref(h))
        if st != 0:
            raise KsolveError(st, f"ks_ctx_create({device}) failed: "
                                  f"{self.lib.ks_status_string(st).decode()}")
        self.h = h
        self.device = device
        self._mats = weakref.WeakSet()  # matrices must be destroyed before the context

    def check(self, st, what):
        if st != 0:
            msg = self.lib.ks_ctx_last_error(self.h).decode(errors="replace")
            raise KsolveError(st, f"{what}: {msg}")

    @property
    def stream(self):
        return self.lib.ks_ctx_stream(self.h)

    def synchronize(self):
        self.check(self.lib.ks_ctx_synchronize(self.h), "synchronize")

    STREAM_KINDS = {"read": 0, "copy": 1}

    def stream_rate(self, kind="read", nbytes=1 << 30, reps=10):
        """GB/s that a plain streaming kernel (read or copy of `nbytes`)
        reaches on this device (ks_bench_stream; measurement only)."""
        out = C.c_double()
        self.check(self.lib.ks_bench_stream(self.h, self.STREAM_KINDS[kind], int(nbytes), int(reps), C.byref(out)),
                   "bench_stream")
        return out.value

    def close(self):
        if getattr(self, "h", None):
            for m in list(self._mats):
                m.close()
            self.lib.ks_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- BLAS layer (blas.cpp) -------------------------------------------
    def dot(self, x, y, precision=64):
        dt = np.float64 if precision == 64 else np.float32
        ax, ay = _Arg(x, dt), _Arg(y, dt)
        out = C.c_double()
        n = ax.obj.numel() if ax.dev else ax.obj.size
        self.check(self.lib.ks_dot(self.h, n, precision, ax.ptr, ay.ptr, C.byref(out), _flags(ax, ay)), "dot")
        return out.value

    def norm2(self, x, precision=64):
        dt = np.float64 if precision == 64 else np.float32
        ax = _Arg(x, dt)
        out = C.c_double()
        n = ax.obj.numel() if ax.dev else ax.obj.size
        self.check(self.lib.ks_norm2(self.h, n, precision, ax.ptr, C.byref(out), _flags(ax)), "norm2")
        return out.value

    def axpy(self, a, x, y):
        ax, ay = _Arg(x, np.float64), _Arg(y, np.float64)
        if ax.dev:
            out = torch.empty_like(ax.obj)
        else:
            out = np.empty_like(ax.obj)
        ao = _Arg(out, np.float64, writable=True)
        n = ax.obj.numel() if ax.dev else ax.obj.size
        self.check(self.lib.ks_axpy(self.h, n, float(a), ax.ptr, ay.ptr, ao.ptr, _flags(ax, ay, ao)), "axpy")
        return out

    def cgs(self, V, w):
        """V: (k+1, n) basis rows, w: (n,).  Returns (w - V^T h, h[0..k+1])."""
        V = np.ascontiguousarray(V, np.float64)
        k = V.shape[0] - 1
        w = np.array(w, np.float64, copy=True)
        h = np.zeros(k + 2, np.float64)
        self.check(self.lib.ks_cgs(self.h, w.size, k, V.ctypes.data, w.ctypes.data, h.ctypes.data, 0), "cgs")
        return w, h

    def kskip_gram(self, k, r, p, Ar, Ap):
        r = np.ascontiguousarray(r, np.float64)
        n = r.size
        d = np.zeros(max(2 * k, 1)); e = np.zeros(2 * k + 1); z = np.zeros(2 * k + 2)
        args = [np.ascontiguousarray(a, np.float64) for a in (p, Ar, Ap)]
        self.check(self.lib.ks_kskip_gram(self.h, n, k, r.ctypes.data, args[0].ctypes.data, args[1].ctypes.data,
                                          args[2].ctypes.data, d.ctypes.data, e.ctypes.data, z.ctypes.data, 0),
                   "kskip_gram")
        return d[:2 * k], e, z


class Matrix:
    """CSR matrix in HBM (collection's Cval/Ccol/Cptr + CTval/... for A^T)."""

    # SpMV layouts: "auto" (the library's choice: column panels for large
    # matrices whose rows gather x from all over, merge-path for heavy-tailed
    # row lengths, CSR-stream on workgroup tiles otherwise), CSR-stream on
    # workgroup ("stream" / "wg") or wavefront tiles, lanes-per-row vector
    # kernel, merge-path, column panels (each panel on the library's choice,
    # or on a named layout: "panels_wg", "panels_wave", "panels_merge");
    # "col32" / "wg_col32": the library's choice / workgroup tiles with int32
    # column indices kept in the SpMV (without it, banded matrices are read
    # with int16 offsets)
    LAYOUTS = {"auto": 0, "col32": _abi.KS_MAT_COL32, "stream": _abi.KS_MAT_WG_TILES, "wg": _abi.KS_MAT_WG_TILES,
               "wg_col32": _abi.KS_MAT_WG_TILES | _abi.KS_MAT_COL32, "delta": _abi.KS_MAT_COL_DELTA,
               "wg_delta": _abi.KS_MAT_WG_TILES | _abi.KS_MAT_COL_DELTA,
               "wave": _abi.KS_MAT_WAVE_TILES, "vector": _abi.KS_MAT_VECTOR, "merge": _abi.KS_MAT_MERGE,
               "panels": _abi.KS_MAT_PANELS, "panels_wg": _abi.KS_MAT_PANELS | _abi.KS_MAT_WG_TILES,
               "panels_wave": _abi.KS_MAT_PANELS | _abi.KS_MAT_WAVE_TILES,
               "panels_merge": _abi.KS_MAT_PANELS | _abi.KS_MAT_MERGE}

    def __init__(self, ctx, ptr, col, val, transpose=False, fp32=False, spd=False, layout="auto",
                 borrow=False):
        self.ctx = ctx
        ap, ac, av = _Arg(ptr, np.int32), _Arg(col, np.int32), _Arg(val, np.float64)
        self.n = (ap.obj.numel() if ap.dev else ap.obj.size) - 1
        self.nnz = ac.obj.numel() if ac.dev else ac.obj.size
        fl = _flags(ap, ac, av)
        if transpose:
            fl |= _abi.KS_MAT_TRANSPOSE
        if fp32:
            fl |= _abi.KS_MAT_FP32_COPY
        if spd:
            fl |= _abi.KS_MAT_SPD
        if layout not in self.LAYOUTS:
            raise ValueError(f"layout must be one of {sorted(self.LAYOUTS)}")
        fl |= self.LAYOUTS[layout]
        if borrow:
            # the library reads these device tensors in place: keep them alive
            fl |= _abi.KS_MAT_BORROW
            self._borrowed = (ap.obj, ac.obj, av.obj)
        h = C.c_void_p()
        ctx.check(ctx.lib.ks_mat_create_csr(ctx.h, self.n, self.nnz, ap.ptr, ac.ptr, av.ptr, fl, C.byref(h)),
                  "ks_mat_create_csr")
        self.h = h
        self.has_t = transpose
        ctx._mats.add(self)

    def info(self):
        n, nnz, kind, lanes = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        self.ctx.check(self.ctx.lib.ks_mat_info(self.h, C.byref(n), C.byref(nnz), C.byref(kind), C.byref(lanes)),
                       "info")
        return dict(n=n.value, nnz=nnz.value, spmv_kind=kind.value, lanes_per_row=lanes.value)

    def index_bytes(self, trans=False, precision=64):
        """Bytes per column index that this matrix's SpMV reads (2: int16
        offsets col - row, 4: int32 columns)."""
        b = C.c_int32()
        self.ctx.check(self.ctx.lib.ks_mat_index_bytes(self.h, int(trans), precision, C.byref(b)), "index_bytes")
        return b.value

    def update_values(self, val):
        av = _Arg(val, np.float64)
        self.ctx.check(self.ctx.lib.ks_mat_update_values(self.h, av.ptr, _flags(av)), "update_values")

    def transpose_csr(self):
        tp = np.empty(self.n + 1, np.int32)
        tc = np.empty(max(self.nnz, 1), np.int32)
        tv = np.empty(max(self.nnz, 1), np.float64)
        self.ctx.check(self.ctx.lib.ks_mat_get_transpose(self.h, tp.ctypes.data, tc.ctypes.data, tv.ctypes.data, 0),
                       "get_transpose")
        return tp, tc[:self.nnz], tv[:self.nnz]

    def spmv(self, x, trans=False, precision=64):
        """y = A x (cuda::MtxVec_mult) or A^T x; FP64 or FP32 values."""
        dt = np.float64 if precision == 64 else np.float32
        ax = _Arg(x, dt)
        y = torch.empty_like(ax.obj) if ax.dev else np.empty(self.n, dt)
        ay = _Arg(y, dt, writable=True)
        self.ctx.check(self.ctx.lib.ks_spmv(self.h, int(trans), precision, ax.ptr, ay.ptr, _flags(ax, ay)), "spmv")
        return y

    def spmv_bench(self, x, y, reps=100, trans=False, precision=64):
        """Mean duration (ms) of `reps` back-to-back SpMV launches (device tensors)."""
        dt = np.float64 if precision == 64 else np.float32
        ax, ay = _Arg(x, dt), _Arg(y, dt, writable=True)
        if not ax.dev:
            raise TypeError("spmv_bench needs device tensors")
        torch.cuda.synchronize()
        ms = C.c_double()
        self.ctx.check(self.ctx.lib.ks_spmv_bench(self.h, int(trans), precision, ax.ptr, ay.ptr, reps, C.byref(ms)),
                       "spmv_bench")
        return ms.value

    def kskip_basis(self, k, r, p):
        r = np.ascontiguousarray(r, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        Ar = np.zeros((2 * k + 1, self.n))
        Ap = np.zeros((2 * k + 2, self.n))
        self.ctx.check(self.ctx.lib.ks_kskip_basis(self.h, k, r.ctypes.data, p.ctypes.data, Ar.ctypes.data,
                                                   Ap.ctypes.data, 0), "kskip_basis")
        return Ar, Ap

    def sddmm(self, alpha, u, v):
        au, av = _Arg(u, np.float64), _Arg(v, np.float64)
        out = torch.empty(self.nnz, dtype=torch.float64, device=au.obj.device) if au.dev \
            else np.empty(self.nnz, np.float64)
        ao = _Arg(out, np.float64, writable=True)
        self.ctx.check(self.ctx.lib.ks_sddmm(self.h, float(alpha), au.ptr, av.ptr, ao.ptr, _flags(au, av, ao)),
                       "sddmm")
        return out

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):  # context still alive
                self.ctx.lib.ks_mat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve(ctx, A, b, outer, inner="NO", x0=None, history=True, hist_cap=1 << 20, **params):
    """Run solver `outer` (with VP inner solver `inner`) on A x = b.

    params use the reference flag names: outer_maxloop (-L), outer_eps (-E),
    outer_restart (-R), outer_kskip (-K), outer_fix (-F), inner_* (-l -e -r
    -k -f), plus inner_precision (64/32), ref_compat (KS_COMPAT_* bits),
    check_every, profile.  Returns (x, result dict, history (m,2) or None).
    """
    p = make_params(outer, inner, **params)
    ab = _Arg(b, np.float64)
    n = A.n
    if ab.dev:
        x = torch.zeros(n, dtype=torch.float64, device=ab.obj.device) if x0 is None else x0.clone()
    else:
        x = np.zeros(n, np.float64) if x0 is None else np.array(x0, np.float64, copy=True)
    ax = _Arg(x, np.float64, writable=True)
    p.flags |= _flags(ab, ax)
    if x0 is not None:
        p.flags |= _abi.KS_X0_GIVEN
    res = KsResult()
    hist = np.zeros((hist_cap, 2), np.float64) if history else None
    ctx.check(ctx.lib.ks_solve(ctx.h, A.h, C.byref(p), ab.ptr, ax.ptr, hist.ctypes.data if history else None,
                               hist_cap if history else 0, C.byref(res)), f"solve({outer}/{inner})")
    d = res.as_dict()
    h = hist[:min(res.n_hist, hist_cap)].copy() if history else None
    return x, d, h


def adjoint(ctx, A, x, g, outer, inner="NO", want_vals=True, **params):
    """Solve A^T lambda = g; returns (grad_b = lambda, grad_vals, result)."""
    p = make_params(outer, inner, **params)
    axx, ag = _Arg(x, np.float64), _Arg(g, np.float64)
    if axx.dev:
        gb = torch.zeros(A.n, dtype=torch.float64, device=axx.obj.device)
        gv = torch.zeros(A.nnz, dtype=torch.float64, device=axx.obj.device) if want_vals else None
    else:
        gb = np.zeros(A.n, np.float64)
        gv = np.zeros(A.nnz, np.float64) if want_vals else None
    agb = _Arg(gb, np.float64, writable=True)
    agv = _Arg(gv, np.float64, writable=True) if want_vals else None
    p.flags |= _flags(axx, ag, agb, agv)
    res = KsResult()
    ctx.check(ctx.lib.ks_adjoint(ctx.h, A.h, C.byref(p), axx.ptr, ag.ptr, agb.ptr,
                                 agv.ptr if want_vals else None, C.byref(res)), f"adjoint({outer})")
    return gb, gv, res.as_dict()
