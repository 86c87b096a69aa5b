"""MI355X (gfx950) Krylov solver library: a drop-in for the Krylov hot path of
ikuno/solver_cpp (CSR input, the reference's solver names and flags), with
hand-written HIP kernels behind a C ABI (include/ksolve.h, libksolve.so).
"""
from . import _abi  # noqa: F401
from ._abi import KS_COMPAT_ALL, KS_COMPAT_KSKIP_AS_WRITTEN, KS_COMPAT_LAST_TERM, KS_COMPAT_VPCG_FR  # noqa: F401

__all__ = ["api", "gen", "autograd"]
