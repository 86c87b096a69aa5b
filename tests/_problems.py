"""Small seeded test problems shared by the CPU and GPU tests."""
import numpy as np

from solver_cpp_amd import gen


def poisson2d(g=32):
    ptr, col, val = gen.stencil(2, g, g)
    xt = gen.vector(1, g * g, "uniform")
    return ptr, col, val, gen.csr_matvec(ptr, col, val, xt), xt


def convdiff3d(g=12, pe=0.5):
    ptr, col, val = gen.stencil(3, g, g, g, pe=pe)
    xt = gen.vector(2, g ** 3, "normal")
    return ptr, col, val, gen.csr_matvec(ptr, col, val, xt), xt


def zipf_small(n=3000, cap=300):
    ptr, col, val = gen.zipf(n, 1.5, cap, (51, 52, 53))
    b = gen.vector(50, n, "uniform")
    return ptr, col, val, b


def aniso(g=24):
    ptr, col, val, kap = gen.aniso2d(g, g, 4)
    return ptr, col, val, kap


def dense(ptr, col, val):
    n = len(ptr) - 1
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(ptr[i], ptr[i + 1]):
            A[i, col[j]] += val[j]
    return A


def row_abs_bound(ptr, col, val, x):
    """sum_j |a_ij x_j| per row (error scale of a row sum)."""
    import scipy.sparse as sp
    n = len(ptr) - 1
    A = sp.csr_matrix((np.abs(val), col, ptr), shape=(n, n))
    return A @ np.abs(x)
