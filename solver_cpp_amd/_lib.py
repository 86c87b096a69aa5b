"""Loads libksolve.so (built in-tree) and declares its C ABI (include/ksolve.h).

There is no fallback: if the library is missing, importing the product API
raises, naming the build command.
"""
import ctypes as C
import os

from ._abi import KsParams, KsResult

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libksolve.so")

_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_u32 = C.c_uint32
_u64 = C.c_uint64
_dbl = C.c_double
_pp = C.POINTER(_vp)

_SIGS = {
    "ks_abi_version": (_i32, []),
    "ks_params_default": (None, [C.POINTER(KsParams)]),
    "ks_status_string": (C.c_char_p, [_i32]),
    "ks_ctx_create": (_i32, [_i32, _pp]),
    "ks_ctx_destroy": (None, [_vp]),
    "ks_ctx_last_error": (C.c_char_p, [_vp]),
    "ks_ctx_stream": (_vp, [_vp]),
    "ks_ctx_synchronize": (_i32, [_vp]),
    "ks_mat_create_csr": (_i32, [_vp, _i64, _i64, _vp, _vp, _vp, _u32, _pp]),
    "ks_mat_update_values": (_i32, [_vp, _vp, _u32]),
    "ks_mat_get_transpose": (_i32, [_vp, _vp, _vp, _vp, _u32]),
    "ks_mat_destroy": (None, [_vp]),
    "ks_mat_info": (_i32, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i32), C.POINTER(_i32)]),
    "ks_mat_index_bytes": (_i32, [_vp, _i32, _i32, C.POINTER(_i32)]),
    "ks_spmv": (_i32, [_vp, _i32, _i32, _vp, _vp, _u32]),
    "ks_spmv_bench": (_i32, [_vp, _i32, _i32, _vp, _vp, _i32, C.POINTER(_dbl)]),
    "ks_bench_stream": (_i32, [_vp, _i32, _i64, _i32, C.POINTER(_dbl)]),
    "ks_dot": (_i32, [_vp, _i64, _i32, _vp, _vp, C.POINTER(_dbl), _u32]),
    "ks_norm2": (_i32, [_vp, _i64, _i32, _vp, C.POINTER(_dbl), _u32]),
    "ks_axpy": (_i32, [_vp, _i64, _dbl, _vp, _vp, _vp, _u32]),
    "ks_cgs": (_i32, [_vp, _i64, _i32, _vp, _vp, _vp, _u32]),
    "ks_kskip_basis": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _u32]),
    "ks_kskip_gram": (_i32, [_vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32]),
    "ks_sddmm": (_i32, [_vp, _dbl, _vp, _vp, _vp, _u32]),
    "ks_solve": (_i32, [_vp, _vp, C.POINTER(KsParams), _vp, _vp, _vp, _i64, C.POINTER(KsResult)]),
    "ks_solve_noalloc": (_i32, [_vp, _vp, C.POINTER(KsParams), _vp, _vp, C.POINTER(KsResult)]),
    "ks_adjoint": (_i32, [_vp, _vp, C.POINTER(KsParams), _vp, _vp, _vp, _vp, C.POINTER(KsResult)]),
    "ks_gen_stencil": (_i32, [_i32, _i32, _i32, _i32, _dbl, _vp, C.POINTER(_i64), C.POINTER(_i64), _vp, _vp, _vp]),
    "ks_gen_aniso2d": (_i32, [_i32, _i32, _u64, C.POINTER(_i64), C.POINTER(_i64), _vp, _vp, _vp, _vp]),
    "ks_aniso2d_values": (_i32, [_i32, _i32, _vp, _vp]),
    "ks_gen_zipf": (_i32, [_i64, _dbl, _i32, _u64, _u64, _u64, C.POINTER(_i64), _vp, _vp, _vp]),
    "ks_gen_vector": (_i32, [_u64, _i32, _dbl, _dbl, _i64, _vp]),
    "ks_crs_read_header": (_i32, [C.c_char_p, C.POINTER(_i64), C.POINTER(_i64)]),
    "ks_crs_read": (_i32, [C.c_char_p, _vp, _vp, _vp, _vp, _vp]),
    "ks_crs_write": (_i32, [C.c_char_p, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "ks_crs_stamp": (_i32, [C.c_char_p, C.POINTER(_u64)]),
    "ks_crsbin_write": (_i32, [C.c_char_p, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _u64]),
    "ks_crsbin_read_header": (_i32, [C.c_char_p, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_u64)]),
    "ks_crsbin_read": (_i32, [C.c_char_p, _vp, _vp, _vp, _vp, _vp]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load():
    """Return the loaded libksolve (raises OSError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}/csrc` "
                          "or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
