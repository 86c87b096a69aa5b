"""ctypes mirrors of the structs and constants in include/ksolve.h.

Pure Python: importing this module loads no shared library.
"""
import ctypes as C

KS_OK = 0
KS_NOT_CONVERGED = 2
KS_BREAKDOWN = 3
KS_ERR_BAD_INPUT = 10
KS_ERR_HIP = 11
KS_ERR_UNSUPPORTED = 12
KS_ERR_NO_MEMORY = 13
KS_ERR_IO = 14

KS_COMPAT_LAST_TERM = 1
KS_COMPAT_KSKIP_AS_WRITTEN = 2
KS_COMPAT_VPCG_FR = 4
KS_COMPAT_ALL = 7

KS_PTR_DEVICE = 1
KS_X0_GIVEN = 2
KS_MAT_FP32_COPY = 4
KS_MAT_TRANSPOSE = 8
KS_MAT_SPD = 16
KS_MAT_MERGE = 32
KS_MAT_VECTOR = 64
KS_MAT_WG_TILES = 128
KS_MAT_BORROW = 256
KS_MAT_WAVE_TILES = 512
KS_MAT_PANELS = 1024
KS_MAT_COL32 = 2048
KS_MAT_COL_DELTA = 4096

KS_F64 = 64
KS_F32 = 32

KERNEL_CLASSES = ("spmv64", "spmv32", "reduce", "cgs", "axpy", "scalar", "basis", "other")
KS_K_NCLASS = len(KERNEL_CLASSES)

STATUS_NAMES = {
    0: "converge", 2: "not converge", 3: "breakdown", 10: "bad input",
    11: "hip error", 12: "unsupported", 13: "no memory", 14: "io error",
}


class KsParams(C.Structure):
    """ks_params: the reference's solver flags (solver_collection.cpp:278-307)."""
    _fields_ = [
        ("outer", C.c_char * 16),
        ("inner", C.c_char * 16),
        ("outer_maxloop", C.c_int32),
        ("outer_restart", C.c_int32),
        ("outer_eps", C.c_double),
        ("outer_kskip", C.c_int32),
        ("outer_fix", C.c_int32),
        ("inner_maxloop", C.c_int32),
        ("inner_restart", C.c_int32),
        ("inner_eps", C.c_double),
        ("inner_kskip", C.c_int32),
        ("inner_fix", C.c_int32),
        ("inner_precision", C.c_int32),
        ("ref_compat", C.c_uint32),
        ("omp_threads", C.c_int32),
        ("verbose", C.c_int32),
        ("flags", C.c_uint32),
        ("check_every", C.c_int32),
        ("profile", C.c_int32),
        ("reserved", C.c_int32),
    ]


class KsResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("reserved", C.c_int32),
        ("iters", C.c_int64),
        ("steps", C.c_int64),
        ("inner_steps", C.c_int64),
        ("spmv", C.c_int64),
        ("n_hist", C.c_int64),
        ("relres", C.c_double),
        ("relres_true", C.c_double),
        ("log10_check", C.c_double),
        ("seconds", C.c_double),
        ("bytes", C.c_double),
        ("prof_ms", C.c_double * KS_K_NCLASS),
        ("prof_launches", C.c_int64 * KS_K_NCLASS),
        ("prof_bytes", C.c_double * KS_K_NCLASS),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_
             if not f.startswith("prof") and f != "reserved"}
        d["prof"] = {
            name: {"ms": self.prof_ms[i], "launches": self.prof_launches[i],
                   "bytes": self.prof_bytes[i]}
            for i, name in enumerate(KERNEL_CLASSES) if self.prof_launches[i]
        }
        return d


# Reference defaults (solver_collection.cpp:20-49).
DEFAULTS = dict(
    outer_maxloop=10000, outer_eps=1e-8, outer_restart=1000, outer_kskip=2, outer_fix=2,
    inner_maxloop=50, inner_eps=1e-1, inner_restart=1000, inner_kskip=2, inner_fix=2,
    inner_precision=KS_F64, ref_compat=0, omp_threads=8, verbose=0, flags=0,
    check_every=0, profile=0,
)


def make_params(outer, inner="NO", **kw):
    """Fill a KsParams from reference-style names (outer/inner solver names and
    the outer_*/inner_* fields); unknown keys raise."""
    p = KsParams()
    vals = dict(DEFAULTS)
    for k, v in kw.items():
        if k not in vals:
            raise KeyError(f"unknown solver parameter {k!r}")
        vals[k] = v
    p.outer = outer.encode()
    p.inner = (inner or "NO").encode()
    for k, v in vals.items():
        setattr(p, k, v)
    return p
