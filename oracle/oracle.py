"""ctypes front end of the CPU oracle (liboracle.so, built from ks_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline,
never as part of the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from solver_cpp_amd._abi import KsParams, KsResult, make_params

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i64 = C.c_int64


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.or_spmv.argtypes = [_i64, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_spmv_f32.argtypes = [_i64, _i32p, _i32p, _f32p, _f32p, _f32p]
        L.or_transpose.argtypes = [_i64, _i64, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p]
        L.or_dot.argtypes = [_i64, _f64p, _f64p, C.c_int32]
        L.or_dot.restype = C.c_double
        L.or_norm2.argtypes = [_i64, _f64p, C.c_int32]
        L.or_norm2.restype = C.c_double
        L.or_cgs.argtypes = [_i64, C.c_int32, _f64p, _f64p, _f64p, C.c_int32]
        L.or_cgs.restype = C.c_double
        L.or_hye.argtypes = [_f64p, C.c_int32, _f64p, _f64p, C.c_int32]
        L.or_kskip_basis.argtypes = [_i64, _i32p, _i32p, _f64p, C.c_int32, _f64p, _f64p, _f64p, _f64p]
        L.or_kskip_gram.argtypes = [_i64, C.c_int32, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.or_sddmm.argtypes = [_i64, _i32p, _i32p, C.c_double, _f64p, _f64p, _f64p]
        L.or_check.argtypes = [_i64, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_void_p,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_solve.argtypes = [C.POINTER(KsParams), _i64, _i64, _i32p, _i32p, _f64p, _f64p, _f64p,
                               C.c_void_p, _i64, C.POINTER(KsResult)]
        L.or_solve.restype = C.c_int32
        L.or_adjoint.argtypes = [C.POINTER(KsParams), _i64, _i64, _i32p, _i32p, _f64p, _f64p, _f64p,
                                 _f64p, _f64p, C.POINTER(KsResult)]
        L.or_adjoint.restype = C.c_int32
        L.or_set_kskip_oob.argtypes = [C.c_double]
        L.or_kskip_trace.argtypes = [C.c_int32]
        L.or_kskip_trace_size.restype = C.c_int64
        L.or_kskip_trace_copy.argtypes = [_f64p]
        L.or_max_threads.restype = C.c_int32
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def spmv(ptr, col, val, x):
    n = len(ptr) - 1
    y = np.empty(n, np.float64)
    lib().or_spmv(n, _c(ptr, np.int32), _c(col, np.int32), _c(val, np.float64), _c(x, np.float64), y)
    return y


def spmv_f32(ptr, col, val, x):
    n = len(ptr) - 1
    y = np.empty(n, np.float32)
    lib().or_spmv_f32(n, _c(ptr, np.int32), _c(col, np.int32), _c(val, np.float32), _c(x, np.float32), y)
    return y


def transpose(ptr, col, val):
    n = len(ptr) - 1
    nnz = len(col)
    tp = np.empty(n + 1, np.int32)
    tc = np.empty(nnz, np.int32)
    tv = np.empty(nnz, np.float64)
    lib().or_transpose(n, nnz, _c(ptr, np.int32), _c(col, np.int32), _c(val, np.float64), tp, tc, tv)
    return tp, tc, tv


def dot(x, y, nchunk=8):
    x = _c(x, np.float64)
    return lib().or_dot(len(x), x, _c(y, np.float64), nchunk)


def norm2(x, nchunk=8):
    x = _c(x, np.float64)
    return lib().or_norm2(len(x), x, nchunk)


def cgs(V, w, k, nchunk=8):
    """V: (k+1, n) rows = basis vectors.  Returns (w_after, h[0..k+1])."""
    V = _c(V, np.float64)
    w = _c(w, np.float64).copy()
    h = np.zeros(k + 2, np.float64)
    lib().or_cgs(len(w), k, V, w, h, nchunk)
    return w, h


def hye(H, e, size):
    H = _c(H, np.float64)
    y = np.zeros(max(size, 1), np.float64)
    lib().or_hye(H, H.shape[1], y, _c(e, np.float64), size)
    return y[:size]


def kskip_basis(ptr, col, val, k, r, p):
    n = len(ptr) - 1
    Ar = np.empty((2 * k + 1, n), np.float64)
    Ap = np.empty((2 * k + 2, n), np.float64)
    lib().or_kskip_basis(n, _c(ptr, np.int32), _c(col, np.int32), _c(val, np.float64), k,
                         _c(r, np.float64), _c(p, np.float64), Ar, Ap)
    return Ar, Ap


def kskip_gram(k, r, p, Ar, Ap):
    n = len(r)
    d = np.empty(2 * k, np.float64)
    e = np.empty(2 * k + 1, np.float64)
    z = np.empty(2 * k + 2, np.float64)
    lib().or_kskip_gram(n, k, _c(r, np.float64), _c(p, np.float64), _c(Ar, np.float64),
                        _c(Ap, np.float64), d if len(d) else np.empty(1), e, z)
    return d, e, z


def sddmm(ptr, col, alpha, u, v):
    n = len(ptr) - 1
    out = np.empty(len(col), np.float64)
    lib().or_sddmm(n, _c(ptr, np.int32), _c(col, np.int32), alpha, _c(u, np.float64), _c(v, np.float64), out)
    return out


def check(ptr, col, val, b, x):
    n = len(ptr) - 1
    lc = C.c_double()
    rr = C.c_double()
    lib().or_check(n, _c(ptr, np.int32), _c(col, np.int32), _c(val, np.float64), _c(b, np.float64),
                   _c(x, np.float64), None, C.byref(lc), C.byref(rr))
    return lc.value, rr.value


def set_kskip_oob(v):
    lib().or_set_kskip_oob(float(v))


def kskip_trace(ptr, col, val, b, k, **kw):
    """Run k-skip CG with the oracle's trace on; returns (x, result, blocks)
    where blocks is a list of (r, p, steps) per block, r and p the block's
    starting vectors and steps a (k+1, 2) array of (alpha, beta)."""
    n = len(ptr) - 1
    L = lib()
    L.or_kskip_trace(1)
    try:
        x, r, _ = solve(ptr, col, val, b, "kskipcg", outer_kskip=k, history=False, **kw)
        buf = np.empty(L.or_kskip_trace_size(), np.float64)
        if len(buf):
            L.or_kskip_trace_copy(buf)
    finally:
        L.or_kskip_trace(0)
    stride = 2 * n + 2 * (k + 1)
    assert len(buf) % stride == 0
    blocks = [(buf[o:o + n], buf[o + n:o + 2 * n], buf[o + 2 * n:o + stride].reshape(k + 1, 2))
              for o in range(0, len(buf), stride)]
    return x, r, blocks


def solve(ptr, col, val, b, outer, inner="NO", x0=None, history=True, hist_cap=1 << 20, **kw):
    """Run the oracle solver; returns (x, result_dict, history (m,2) array)."""
    n = len(ptr) - 1
    p = make_params(outer, inner, **kw)
    x = np.zeros(n, np.float64) if x0 is None else _c(x0, np.float64).copy()
    if x0 is not None:
        p.flags |= 2
    res = KsResult()
    hist = np.zeros((hist_cap, 2), np.float64) if history else None
    st = lib().or_solve(C.byref(p), n, len(col), _c(ptr, np.int32), _c(col, np.int32),
                        _c(val, np.float64), _c(b, np.float64), x,
                        hist.ctypes.data if history else None, hist_cap if history else 0,
                        C.byref(res))
    if st not in (0, 2):
        raise RuntimeError(f"oracle solve failed with status {st}")
    d = res.as_dict()
    h = hist[:min(res.n_hist, hist_cap)].copy() if history else None
    return x, d, h


def adjoint(ptr, col, val, x, g, outer, inner="NO", **kw):
    n = len(ptr) - 1
    p = make_params(outer, inner, **kw)
    res = KsResult()
    gb = np.zeros(n, np.float64)
    gv = np.zeros(len(col), np.float64)
    st = lib().or_adjoint(C.byref(p), n, len(col), _c(ptr, np.int32), _c(col, np.int32), _c(val, np.float64),
                          _c(x, np.float64), _c(g, np.float64), gb, gv, C.byref(res))
    if st not in (0, 2):
        raise RuntimeError(f"oracle adjoint failed with status {st}")
    return gb, gv, res.as_dict()


def max_threads():
    return lib().or_max_threads()
