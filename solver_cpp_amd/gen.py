"""Deterministic synthetic inputs (host) through libksolve's generators.

The reference reads matrices from CRS text files and ships none; these
generators build the benchmark problems (5/7-point stencils, the log-normal
anisotropic diffusion and the Zipf random matrix) with a documented RNG
(xoshiro256** seeded by splitmix64, Box-Muller normals) so every run sees
identical inputs.
"""
import ctypes as C
import os

import numpy as np

from ._lib import load


def _chk(st, what):
    if st != 0:
        raise RuntimeError(f"{what} failed with status {st}")


def stencil(dim, nx, ny, nz=1, pe=0.0, vel=(1.0, 0.5, 0.25)):
    """5-point (dim 2) / 7-point (dim 3) upwind convection-diffusion CSR."""
    L = load()
    n, nnz = C.c_int64(), C.c_int64()
    v = np.ascontiguousarray(vel, np.float64)
    _chk(L.ks_gen_stencil(dim, nx, ny, nz, pe, v.ctypes.data, C.byref(n), C.byref(nnz), None, None, None),
         "ks_gen_stencil")
    ptr = np.empty(n.value + 1, np.int32)
    col = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    _chk(L.ks_gen_stencil(dim, nx, ny, nz, pe, v.ctypes.data, None, None, ptr.ctypes.data, col.ctypes.data,
                          val.ctypes.data), "ks_gen_stencil")
    return ptr, col, val


def aniso2d(nx, ny, seed):
    """2-D graph Laplacian with edge conductivities exp(N(0,1)) (SPD)."""
    L = load()
    n, nnz = C.c_int64(), C.c_int64()
    _chk(L.ks_gen_aniso2d(nx, ny, seed, C.byref(n), C.byref(nnz), None, None, None, None), "ks_gen_aniso2d")
    ne = (nx + 1) * ny + nx * (ny + 1)
    ptr = np.empty(n.value + 1, np.int32)
    col = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    kap = np.empty(ne, np.float64)
    _chk(L.ks_gen_aniso2d(nx, ny, seed, None, None, ptr.ctypes.data, col.ctypes.data, val.ctypes.data,
                          kap.ctypes.data), "ks_gen_aniso2d")
    return ptr, col, val, kap


def aniso2d_values(nx, ny, kappa):
    """CSR values of the aniso2d matrix for edge conductivities `kappa` (linear in kappa)."""
    L = load()
    kappa = np.ascontiguousarray(kappa, np.float64)
    n = nx * ny
    nnz = n + 2 * ((nx - 1) * ny + nx * (ny - 1))
    val = np.empty(nnz, np.float64)
    _chk(L.ks_aniso2d_values(nx, ny, kappa.ctypes.data, val.ctypes.data), "ks_aniso2d_values")
    return val


def zipf(n, s=1.5, cap=4096, seeds=(51, 52, 53)):
    """Random nonsymmetric diagonally dominant matrix with Zipf(s) row lengths."""
    L = load()
    nnz = C.c_int64()
    _chk(L.ks_gen_zipf(n, s, cap, seeds[0], seeds[1], seeds[2], C.byref(nnz), None, None, None), "ks_gen_zipf")
    ptr = np.empty(n + 1, np.int32)
    col = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    _chk(L.ks_gen_zipf(n, s, cap, seeds[0], seeds[1], seeds[2], None, ptr.ctypes.data, col.ctypes.data,
                       val.ctypes.data), "ks_gen_zipf")
    return ptr, col, val


def vector(seed, n, dist="uniform", lo=-1.0, hi=1.0):
    L = load()
    out = np.empty(n, np.float64)
    _chk(L.ks_gen_vector(seed, 0 if dist == "uniform" else 1, lo, hi, n, out.ctypes.data), "ks_gen_vector")
    return out


def csr_matvec(ptr, col, val, x):
    """Host y = A x with sequential row sums (for building b = A x_true)."""
    import scipy.sparse as sp
    n = len(ptr) - 1
    A = sp.csr_matrix((val, col, ptr), shape=(n, n))
    return A @ x


def write_crs(path, ptr, col, val, b, x=None):
    """Write the reference's CRS text files (ColVal.txt, Ptr.txt, bx.txt)."""
    os.makedirs(path, exist_ok=True)
    L = load()
    n = len(ptr) - 1
    ptr = np.ascontiguousarray(ptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n) if x is None else np.ascontiguousarray(x, np.float64)
    _chk(L.ks_crs_write(path.encode(), n, len(col), ptr.ctypes.data, col.ctypes.data, val.ctypes.data,
                        b.ctypes.data, x.ctypes.data), "ks_crs_write")


def read_crs(path):
    L = load()
    n, nnz = C.c_int64(), C.c_int64()
    _chk(L.ks_crs_read_header(path.encode(), C.byref(n), C.byref(nnz)), "ks_crs_read_header")
    ptr = np.empty(n.value + 1, np.int32)
    col = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    b = np.empty(n.value, np.float64)
    x = np.empty(n.value, np.float64)
    _chk(L.ks_crs_read(path.encode(), ptr.ctypes.data, col.ctypes.data, val.ctypes.data, b.ctypes.data,
                       x.ctypes.data), "ks_crs_read")
    return ptr, col, val, b, x


def crs_stamp(path):
    """Fingerprint of a CRS text directory (sizes, modification times)."""
    st = C.c_uint64()
    _chk(load().ks_crs_stamp(path.encode(), C.byref(st)), "ks_crs_stamp")
    return st.value


def write_crs_bin(file, ptr, col, val, b=None, x=None, stamp=0):
    """Write the binary CSR sidecar (ks_crsbin_write)."""
    n = len(ptr) - 1
    ptr = np.ascontiguousarray(ptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    bp = None if b is None else np.ascontiguousarray(b, np.float64)
    xp = None if x is None else np.ascontiguousarray(x, np.float64)
    _chk(load().ks_crsbin_write(file.encode(), n, len(col), ptr.ctypes.data, col.ctypes.data, val.ctypes.data,
                                None if bp is None else bp.ctypes.data, None if xp is None else xp.ctypes.data,
                                stamp), "ks_crsbin_write")


def read_crs_bin(file):
    """Read the binary CSR sidecar: (ptr, col, val, b, x, stamp)."""
    L = load()
    n, nnz, st = C.c_int64(), C.c_int64(), C.c_uint64()
    _chk(L.ks_crsbin_read_header(file.encode(), C.byref(n), C.byref(nnz), C.byref(st)), "ks_crsbin_read_header")
    ptr = np.empty(n.value + 1, np.int32)
    col = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    b = np.empty(n.value, np.float64)
    x = np.empty(n.value, np.float64)
    _chk(L.ks_crsbin_read(file.encode(), ptr.ctypes.data, col.ctypes.data, val.ctypes.data, b.ctypes.data,
                          x.ctypes.data), "ks_crsbin_read")
    return ptr, col, val, b, x, st.value


# --- benchmark problems (seeds fixed by the benchmark plan) ---------------
def problem(cfg, scale=None):
    """Matrix and right-hand side of a benchmark configuration.

    cfg: "poisson2d" (256^2, x_true~U(-1,1) seed 1, b=A x_true),
         "convdiff3d" (128^3 upwind Pe 0.5, x_true~N(0,1) seed 2),
         "poisson3d_sym" (128^3 Pe 0 variant, same x_true),
         "poisson3d" (160^3, b~U(-1,1) seed 3),
         "aniso2d" (1024^2 log-normal edges seed 4, b~U(-1,1) seed 41),
         "convdiff3d_pe08" (192^3 Pe 0.8, b~U(-1,1) seed 50),
         "zipf" (n=4M, seeds 51/52/53, b~U(-1,1) seed 50).
    scale overrides the grid edge (or n for zipf) for small test cases.
    """
    if cfg == "poisson2d":
        g = scale or 256
        ptr, col, val = stencil(2, g, g)
        xt = vector(1, g * g, "uniform")
        return dict(ptr=ptr, col=col, val=val, b=csr_matvec(ptr, col, val, xt), x_true=xt)
    if cfg in ("convdiff3d", "poisson3d_sym"):
        g = scale or 128
        ptr, col, val = stencil(3, g, g, g, pe=0.5 if cfg == "convdiff3d" else 0.0)
        xt = vector(2, g ** 3, "normal")
        return dict(ptr=ptr, col=col, val=val, b=csr_matvec(ptr, col, val, xt), x_true=xt)
    if cfg == "poisson3d":
        g = scale or 160
        ptr, col, val = stencil(3, g, g, g)
        return dict(ptr=ptr, col=col, val=val, b=vector(3, g ** 3, "uniform"), x_true=None)
    if cfg == "aniso2d":
        g = scale or 1024
        ptr, col, val, kap = aniso2d(g, g, 4)
        return dict(ptr=ptr, col=col, val=val, b=vector(41, g * g, "uniform"), x_true=None, kappa=kap,
                    x_target=vector(40, g * g, "normal"), grid=g)
    if cfg == "convdiff3d_pe08":
        g = scale or 192
        ptr, col, val = stencil(3, g, g, g, pe=0.8)
        return dict(ptr=ptr, col=col, val=val, b=vector(50, g ** 3, "uniform"), x_true=None,
                    x_target=vector(5, g ** 3, "normal"))
    if cfg == "zipf":
        n = scale or 4_000_000
        ptr, col, val = zipf(n, 1.5, 4096, (51, 52, 53))
        return dict(ptr=ptr, col=col, val=val, b=vector(50, n, "uniform"), x_true=None,
                    x_target=vector(5, n, "normal"))
    raise KeyError(cfg)
