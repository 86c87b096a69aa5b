"""Pins of the CPU oracle (oracle/ks_oracle.cpp) - CPU only.

The reference ships no tests, fixtures or golden vectors (its repository has
none), and its host path cannot be built here (it needs the CUDA runtime and
cuSPARSE), so each oracle function is pinned by what the mathematics fixes:

  function            pinned by
  ------------------  ---------------------------------------------------------
  spmv / spmv_f32     exact sums (math.fsum) within the row-sum rounding
                      bound; A*1 closed form; pure sequential order
  transpose           (A^T)^T = A, dense transpose, y^T A x = x^T A^T y,
                      column counts, rows sorted; with repeated (row, col)
                      entries, numpy's stable argsort by column
  dot / norm2         sum_{i<n} i = n(n-1)/2; math.fsum bound
  cgs                 orthogonality, h = V w, ||w_out|| = h[k+1]
  hye                 triangular solve (numpy)
  kskip_basis / gram  repeated products A^{i+1} r; numpy dot products
  sddmm               dense outer product on the pattern
  cg                  1-D Poisson b=1 closed form x_i = i(n+1-i)/2; a
                      diagonal A with k distinct eigenvalues stops after k
                      iterations; scipy spsolve solution
  bicg                = CG on SPD matrices (same history); spsolve
  gmres               min-residual history = least squares over the Krylov
                      space (brute force); Givens estimate = true residual;
                      KS_COMPAT_LAST_TERM leaves the history unchanged, and
                      its x - x0 = y_{m-1} v_{m-1} of a numpy Arnoldi +
                      lstsq (the textbook x - x0 = V_m y)
  kskipcg (derived)   = CG history at the block starts (exact arithmetic);
                      every step of a block = a textbook CG step (numpy)
                      from the block's (r, p); k = 0 = CG
  kskipcg (as written)k = 0 identical to the derived form (no update runs);
                      the first step of every block = a textbook CG step
                      from the block's (r, p); invariant to the out-of-range
                      eta read (0, 1e300, NaN).  The as-written delta/eta/
                      zeta updates feeding steps 2..k+1 of a block: parity
                      unpinned (no independent source)
  kskipbicg           = BiCG history at the block starts
  cr, gcr             spsolve solution; GCR(restart) = GMRES history
  vpgmres             inner = one CG step (z = c v) reproduces GMRES;
                      FP32 inner still reaches the FP64 tolerance; the
                      inner GMRES preconditioner output (last Krylov term
                      with KS_COMPAT_LAST_TERM, else V_m y) from numpy
  vpcg                inner = one CG step reproduces CG (both beta forms)
  adjoint             diagonal closed form; dense solve; central FD
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from solver_cpp_amd import _abi
from tests import _problems as P

EPS = 2.0 ** -53


def _csr(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def poisson1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")


# ----------------------------------------------------------------- BLAS ----
def test_spmv_exact_sums_and_order(oracle):
    ptr, col, val, _ = P.zipf_small(3000, 200)
    x = np.random.default_rng(0).standard_normal(3000)
    y = oracle.spmv(ptr, col, val, x)
    bound = P.row_abs_bound(ptr, col, val, x)
    for i in range(0, 3000, 7):
        s = slice(ptr[i], ptr[i + 1])
        exact = math.fsum(val[s] * x[col[s]])
        m = ptr[i + 1] - ptr[i]
        assert abs(y[i] - exact) <= m * EPS * bound[i] + 1e-300
        # sequential CSR-order sum, products rounded (blas.cpp:56-62)
        t = 0.0
        for j in range(ptr[i], ptr[i + 1]):
            t += float(val[j]) * float(x[col[j]])
        assert y[i] == t
    # closed form: Poisson A*1 is 0 inside and the count of missing neighbours on the boundary
    ptr, col, val, _, _ = P.poisson2d(9)
    y = oracle.spmv(ptr, col, val, np.ones(81))
    deg = np.diff(ptr) - 1
    assert np.array_equal(y, 4.0 - deg)
    y32 = oracle.spmv_f32(ptr, col, val.astype(np.float32), np.ones(81, np.float32))
    assert np.array_equal(y32, (4.0 - deg).astype(np.float32))


def test_transpose(oracle):
    ptr, col, val, _ = P.zipf_small(700, 80)
    tp, tc, tv = oracle.transpose(ptr, col, val)
    D = P.dense(ptr, col, val)
    assert np.array_equal(P.dense(tp, tc, tv), D.T)
    assert np.array_equal(np.diff(tp), np.bincount(col, minlength=700))
    for i in range(700):
        assert np.all(np.diff(tc[tp[i]:tp[i + 1]]) > 0)
    p2, c2, v2 = oracle.transpose(tp, tc, tv)
    assert np.array_equal(p2, ptr) and np.array_equal(c2, col) and np.array_equal(v2, val)
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(700), rng.standard_normal(700)
    assert abs(y @ oracle.spmv(ptr, col, val, x) - x @ oracle.spmv(tp, tc, tv, y)) < 1e-10


def test_transpose_repeated_entries(oracle):
    # unsorted rows with repeated (row, column) entries: A^T keeps the entries
    # of each column in CSR order, i.e. a stable sort of the entries by column
    rng = np.random.default_rng(4)
    n = 300
    rl = rng.integers(0, 9, n)
    ptr = np.zeros(n + 1, np.int32)
    ptr[1:] = np.cumsum(rl)
    col = rng.integers(0, 40, int(ptr[-1])).astype(np.int32)  # many repeats
    val = rng.standard_normal(len(col))
    tp, tc, tv = oracle.transpose(ptr, col, val)
    order = np.argsort(col, kind="stable")
    rows = np.repeat(np.arange(n, dtype=np.int32), rl)
    assert np.array_equal(tp, np.concatenate([[0], np.cumsum(np.bincount(col, minlength=n))]))
    assert np.array_equal(tc, rows[order]) and np.array_equal(tv, val[order])


@pytest.mark.parametrize("n", [1, 10, 12345, 1_000_003])
def test_dot_norm(oracle, n):
    i = np.arange(n, dtype=np.float64)
    assert oracle.dot(i, np.ones(n)) == n * (n - 1) / 2
    x = np.random.default_rng(n).standard_normal(n)
    y = np.random.default_rng(n + 1).standard_normal(n)
    for nc in (1, 8, 16):
        assert abs(oracle.dot(x, y, nc) - math.fsum(x * y)) <= n * EPS * np.abs(x * y).sum() + 1e-300
    assert abs(oracle.norm2(x) - math.sqrt(math.fsum(x * x))) <= n * EPS * np.linalg.norm(x)


def test_cgs_and_hye(oracle):
    rng = np.random.default_rng(2)
    n, k = 2000, 6
    V = np.linalg.qr(rng.standard_normal((n, k + 1)))[0].T.copy()
    c = rng.standard_normal(k + 1)
    perp = rng.standard_normal(n)
    perp -= V.T @ (V @ perp)
    w = V.T @ c + perp
    wo, h = oracle.cgs(V, w, k)
    assert np.allclose(h[:k + 1], c, atol=1e-12)
    assert np.allclose(h[:k + 1], V @ w, atol=1e-12)
    assert abs(h[k + 1] - np.linalg.norm(wo)) <= 1e-12 * h[k + 1]
    assert np.abs(V @ wo).max() < 1e-12
    assert np.allclose(wo, perp, atol=1e-12)
    # back substitution (blas::Hye) against a dense triangular solve
    R = np.triu(rng.standard_normal((8, 8))) + 4 * np.eye(8)
    e = rng.standard_normal(8)
    for size in (1, 5, 8):
        y = oracle.hye(R, e, size)
        assert np.allclose(y, np.linalg.solve(R[:size, :size], e[:size]), rtol=1e-12)


@pytest.mark.parametrize("k", [1, 2, 4])
def test_kskip_basis_gram(oracle, k):
    ptr, col, val, _, _ = P.convdiff3d(7, pe=0.0)
    n = len(ptr) - 1
    A = sp.csr_matrix((val, col, ptr), shape=(n, n))
    rng = np.random.default_rng(k)
    r, p = rng.standard_normal(n), rng.standard_normal(n)
    Ar, Ap = oracle.kskip_basis(ptr, col, val, k, r, p)
    assert np.array_equal(Ar[0], oracle.spmv(ptr, col, val, r))
    vr, vp = r.copy(), p.copy()
    for i in range(2 * k + 2):
        vr, vp = A @ vr, A @ vp
        if i < 2 * k + 1:
            assert np.allclose(Ar[i], vr, rtol=1e-12, atol=1e-12 * np.abs(vr).max())
        assert np.allclose(Ap[i], vp, rtol=1e-12, atol=1e-12 * np.abs(vp).max())
    d, e, z = oracle.kskip_gram(k, r, p, Ar, Ap)
    assert np.allclose(d, Ar[:2 * k] @ r, rtol=1e-12)
    assert np.allclose(e, Ap[:2 * k + 1] @ r, rtol=1e-12)
    assert np.allclose(z, Ap @ p, rtol=1e-12)


def test_sddmm(oracle):
    ptr, col, val, _ = P.zipf_small(300, 30)
    rng = np.random.default_rng(4)
    u, v = rng.standard_normal(300), rng.standard_normal(300)
    out = oracle.sddmm(ptr, col, -2.0, u, v)
    rows = np.repeat(np.arange(300), np.diff(ptr))
    assert np.array_equal(out, -2.0 * u[rows] * v[col])


# -------------------------------------------------------------- solvers ----
def test_cg_poisson1d_closed_form(oracle):
    n = 200
    ptr, col, val = _csr(poisson1d(n))
    x, r, h = oracle.solve(ptr, col, val, np.ones(n), "cg", outer_eps=1e-13, outer_maxloop=1000)
    i = np.arange(1, n + 1, dtype=np.float64)
    exact = i * (n + 1 - i) / 2
    assert r["status"] == 0
    assert np.linalg.norm(x - exact) <= 1e-9 * np.linalg.norm(exact)
    # CG on an n x n tridiagonal SPD matrix with b = 1: at most n/2 steps (symmetric spectrum pairs)
    assert r["iters"] <= n // 2 + 2


def test_cg_distinct_eigenvalues(oracle):
    # diag with 5 distinct values: CG terminates after 5 iterations, so the
    # test at loop 6 passes (cg.cpp:184-198 reports the passing loop index)
    d = np.repeat([1.0, 2.0, 3.5, 7.0, 11.0], 40)
    ptr, col, val = _csr(sp.diags(d))
    b = np.random.default_rng(3).standard_normal(len(d))
    x, r, h = oracle.solve(ptr, col, val, b, "cg", outer_eps=1e-12, outer_maxloop=100)
    assert r["status"] == 0 and r["iters"] == 6
    assert np.allclose(x, b / d, rtol=1e-12)
    assert h[0, 0] == 1 and h[0, 1] == 1.0


@pytest.mark.parametrize("name,kw,spd", [
    ("cg", {}, True), ("bicg", {}, False), ("cr", {}, True), ("gmres", dict(outer_restart=30), False),
    ("gcr", dict(outer_restart=30), False), ("kskipcg", dict(outer_kskip=3), True),
    ("kskipbicg", dict(outer_kskip=3), False)])
def test_solvers_match_spsolve(oracle, name, kw, spd):
    ptr, col, val, b, xt = P.convdiff3d(9, pe=0.0 if spd else 0.5)
    n = len(ptr) - 1
    x, r, h = oracle.solve(ptr, col, val, b, name, outer_eps=1e-11, outer_maxloop=4000, **kw)
    xs = spla.spsolve(sp.csc_matrix(sp.csr_matrix((val, col, ptr), shape=(n, n))), b)
    assert r["status"] == 0
    assert r["relres_true"] <= 2e-11
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)
    assert np.allclose(x, xt, rtol=0, atol=1e-8 * np.abs(xt).max())


def test_bicg_equals_cg_on_spd(oracle):
    ptr, col, val, b, _ = P.poisson2d(20)
    _, rc, hc = oracle.solve(ptr, col, val, b, "cg", outer_eps=1e-10)
    _, rb, hb = oracle.solve(ptr, col, val, b, "bicg", outer_eps=1e-10)
    assert rc["iters"] == rb["iters"]
    m = min(len(hc), len(hb))
    assert np.allclose(hc[:m], hb[:m], rtol=1e-8)


def _krylov_min_residuals(A, b, kmax):
    """min over x in K_k of ||b - A x|| / ||b||, k = 0..kmax (brute force)."""
    out = [1.0]
    K = [b / np.linalg.norm(b)]
    for k in range(1, kmax + 1):
        Q = np.linalg.qr(np.array(K).T)[0]
        AQ = A @ Q
        y = np.linalg.lstsq(AQ, b, rcond=None)[0]
        out.append(np.linalg.norm(b - AQ @ y) / np.linalg.norm(b))
        K.append(A @ K[-1])
        K[-1] /= np.linalg.norm(K[-1])
    return np.array(out)


def test_gmres_minimal_residual_history(oracle):
    ptr, col, val, b, _ = P.convdiff3d(5)
    D = P.dense(ptr, col, val)
    x, r, h = oracle.solve(ptr, col, val, b, "gmres", outer_eps=1e-12, outer_restart=50, outer_maxloop=40)
    ref = _krylov_min_residuals(D, b, 8)
    assert np.allclose(h[:9, 1], ref, rtol=1e-7)
    assert np.array_equal(h[:, 0], np.arange(1, len(h) + 1))
    # converged inside the first cycle: Givens estimate = true residual
    assert r["status"] == 0 and abs(r["relres"] - r["relres_true"]) <= 1e-3 * r["relres_true"]


def test_gmres_last_term_compat(oracle):
    # Scalar_ax assigns (blas.cpp:211-224): only the assembly of x changes
    ptr, col, val, b, _ = P.convdiff3d(8)
    kw = dict(outer_eps=1e-10, outer_restart=200, outer_maxloop=500)
    x0, r0, h0 = oracle.solve(ptr, col, val, b, "gmres", **kw)
    x1, r1, h1 = oracle.solve(ptr, col, val, b, "gmres", ref_compat=_abi.KS_COMPAT_LAST_TERM, **kw)
    assert r0["iters"] == r1["iters"] and np.array_equal(h0, h1)
    assert r0["relres_true"] < 2e-10 < r1["relres_true"]


def test_gmres_restart_semantics(oracle):
    # -R R runs R-1 Arnoldi steps per cycle (gmres.cpp:277); the limit trips at count+1 >= maxloop
    ptr, col, val, b, _ = P.convdiff3d(8)
    x, r, h = oracle.solve(ptr, col, val, b, "gmres", outer_eps=1e-15, outer_restart=5, outer_maxloop=10)
    assert r["status"] == 2 and r["iters"] == 10 and r["steps"] == 9
    assert np.array_equal(h[:, 0], np.arange(1, 11))


def test_gcr_equals_gmres(oracle):
    ptr, col, val, b, _ = P.convdiff3d(7)
    _, rg, hg = oracle.solve(ptr, col, val, b, "gmres", outer_eps=1e-10, outer_restart=400, outer_maxloop=300)
    _, rc, hc = oracle.solve(ptr, col, val, b, "gcr", outer_eps=1e-10, outer_restart=300, outer_maxloop=300)
    m = min(len(hg), len(hc), 30)
    assert np.allclose(hg[:m, 1], hc[:m, 1], rtol=1e-6)


@pytest.mark.parametrize("k", [1, 2, 4])
def test_kskipcg_derived_equals_cg(oracle, k):
    ptr, col, val, b, _ = P.convdiff3d(8, pe=0.0)
    _, rc, hc = oracle.solve(ptr, col, val, b, "cg", outer_eps=1e-10)
    _, rk, hk = oracle.solve(ptr, col, val, b, "kskipcg", outer_eps=1e-10, outer_kskip=k)
    cg_at = dict(zip(hc[:, 0].astype(int), hc[:, 1]))
    for idx, v in hk[:6]:
        assert abs(v - cg_at[int(idx)]) <= 1e-6 * cg_at[int(idx)]
    assert rk["status"] == 0


def test_kskipcg_as_written_oob_invariance(oracle):
    # j = 2k-1 reads eta[2k+1] past the array (kskipcg.cpp:284); the value
    # read must not reach alpha/beta
    ptr, col, val, b, _ = P.convdiff3d(7, pe=0.0)
    kw = dict(outer_eps=1e-10, outer_kskip=3, outer_maxloop=60, ref_compat=_abi.KS_COMPAT_KSKIP_AS_WRITTEN)
    outs = []
    for v in (0.0, 1e300, float("nan"), -7.0):
        oracle.set_kskip_oob(v)
        outs.append(oracle.solve(ptr, col, val, b, "kskipcg", **kw))
    oracle.set_kskip_oob(0.0)
    for x, r, h in outs[1:]:
        assert np.array_equal(h, outs[0][2]) and np.array_equal(x, outs[0][0])


def _cg_steps(A, r, p, m):
    """m textbook CG steps from (r, p): [(alpha, beta)], and the final (r, p)."""
    out = []
    for _ in range(m):
        Ap = A @ p
        rr = r @ r
        alpha = rr / (p @ Ap)
        r = r - alpha * Ap
        beta = (r @ r) / rr
        p = r + beta * p
        out.append((alpha, beta))
    return np.array(out), r, p


def _spd_case():
    ptr, col, val, b, _ = P.convdiff3d(6, pe=0.0)
    return ptr, col, val, b, sp.csr_matrix((val, col, ptr), shape=(len(ptr) - 1,) * 2)


@pytest.mark.parametrize("compat", [0, _abi.KS_COMPAT_KSKIP_AS_WRITTEN])
def test_kskipcg_k0_is_cg(oracle, compat):
    # k = 0: no delta/eta/zeta update runs (the loop bound 2k - 2(i - nloop)
    # is 0, kskipcg.cpp:283), so the as-written and derived forms are the same
    # computation, and each block is one CG step with beta = alpha zeta_1 /
    # zeta_0 - 1 = (r', r') / (r, r)
    ptr, col, val, b, A = _spd_case()
    kw = dict(outer_eps=1e-10, outer_maxloop=500)
    xk, rk, hk = oracle.solve(ptr, col, val, b, "kskipcg", outer_kskip=0, ref_compat=compat, **kw)
    xd, rd, hd = oracle.solve(ptr, col, val, b, "kskipcg", outer_kskip=0, **kw)
    assert np.array_equal(xk, xd) and np.array_equal(hk, hd)
    xc, rc, hc = oracle.solve(ptr, col, val, b, "cg", **kw)
    assert rk["status"] == rc["status"] == 0 and abs(rk["iters"] - rc["iters"]) <= 1
    assert np.linalg.norm(xk - xc) <= 1e-10 * np.linalg.norm(xc)
    assert np.array_equal(hk[:, 0], hc[:len(hk), 0])


@pytest.mark.parametrize("compat", [0, _abi.KS_COMPAT_KSKIP_AS_WRITTEN])
def test_kskipcg_block_start_steps_are_cg_steps(oracle, compat):
    # the first inner step of every block uses only the Gram products of that
    # block's (r, p) (kskipcg.cpp:262-271): alpha = (r,r)/(p,Ap) and beta =
    # alpha (Ap,Ap)/(p,Ap) - 1 (zeta_1 = (p, A^2 p) = (Ap, Ap) for symmetric
    # A), from inner products numpy forms on the traced vectors - a wrong Gram
    # index or product fails here, in both forms.  (This beta is the textbook
    # (r',r')/(r,r) while p stays A-conjugate to the previous direction, which
    # the derived form keeps and the diverging as-written form loses.)
    ptr, col, val, b, A = _spd_case()
    k = 3
    x, r, blocks = oracle.kskip_trace(ptr, col, val, b, k, outer_eps=1e-10, outer_maxloop=40, ref_compat=compat)
    assert len(blocks) >= 5
    for rb, pb, steps in blocks:
        Ap = A @ pb
        alpha = (rb @ rb) / (pb @ Ap)
        beta = alpha * (Ap @ Ap) / (pb @ Ap) - 1.0
        assert abs(steps[0, 0] - alpha) <= 1e-10 * abs(alpha)
        assert abs(steps[0, 1] - beta) <= 1e-9 * max(1.0, abs(beta))
        if not compat:
            ref, _, _ = _cg_steps(A, rb, pb, 1)
            assert abs(steps[0, 1] - ref[0, 1]) <= 1e-8 * max(1.0, abs(ref[0, 1]))


def test_kskipcg_derived_every_step_is_cg(oracle):
    # derived recurrences: all k+1 steps of a block are textbook CG steps from
    # the block's (r, p), and the next block starts from the r, p they reach
    ptr, col, val, b, A = _spd_case()
    k = 3
    x, r, blocks = oracle.kskip_trace(ptr, col, val, b, k, outer_eps=1e-10, outer_maxloop=40)
    for i, (rb, pb, steps) in enumerate(blocks):
        ref, r1, p1 = _cg_steps(A, rb, pb, k + 1)
        assert np.allclose(steps, ref, rtol=1e-7, atol=1e-12), (i, steps, ref)
        if i + 1 < len(blocks):
            rn, pn, _ = blocks[i + 1]
            assert np.linalg.norm(rn - r1) <= 1e-8 * np.linalg.norm(rb)
            assert np.linalg.norm(pn - p1) <= 1e-8 * np.linalg.norm(pb)
    # as written, the later steps of a block leave textbook CG: the switch is live
    _, _, bw = oracle.kskip_trace(ptr, col, val, b, k, outer_eps=1e-10, outer_maxloop=8,
                                  ref_compat=_abi.KS_COMPAT_KSKIP_AS_WRITTEN)
    ref, _, _ = _cg_steps(A, bw[0][0], bw[0][1], k + 1)
    assert np.max(np.abs(bw[0][2][1:, 0] - ref[1:, 0]) / np.abs(ref[1:, 0])) > 1e-6


def _arnoldi(A, v, m):
    """m Arnoldi steps (classical Gram-Schmidt) from v: V (n, m+1), Hbar (m+1, m)."""
    n = len(v)
    V = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    V[:, 0] = v / np.linalg.norm(v)
    for k in range(m):
        w = A @ V[:, k]
        h = V[:, :k + 1].T @ w
        w = w - V[:, :k + 1] @ h
        H[:k + 1, k] = h
        H[k + 1, k] = np.linalg.norm(w)
        V[:, k + 1] = w / H[k + 1, k]
    return V, H


def _gmres_parts(A, b, m):
    """GMRES after m steps from x0 = 0: (V, y) with y the least-squares
    solution of Hbar y = ||b|| e_1 (numpy lstsq)."""
    V, H = _arnoldi(A, b, m)
    e = np.zeros(m + 1)
    e[0] = np.linalg.norm(b)
    y = np.linalg.lstsq(H, e, rcond=None)[0]
    return V, y


def _tiny_nonsym(n=14, seed=3):
    rng = np.random.default_rng(seed)
    D = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.4)
    D += np.diag(np.abs(D).sum(1) + 1.0)
    ptr, col, val = _csr(D)
    return ptr, col, val, rng.standard_normal(n), D


def test_gmres_last_term_x_is_last_krylov_term(oracle):
    # the limit exit after m Arnoldi steps (count+1 >= maxloop, gmres.cpp:286)
    # assembles x with Scalar_ax, which assigns (blas.cpp:211-224): x = x0 +
    # y_{m-1} v_{m-1} from the m-step least-squares y; the textbook sum is
    # x0 + V_m y.  Both from an independent numpy Arnoldi + lstsq.
    ptr, col, val, b, D = _tiny_nonsym()
    m = 6
    V, y = _gmres_parts(D, b, m)
    last, full = y[m - 1] * V[:, m - 1], V[:, :m] @ y
    kw = dict(outer_eps=0.0, outer_restart=50, outer_maxloop=m + 1)
    xc, rc, _ = oracle.solve(ptr, col, val, b, "gmres", ref_compat=_abi.KS_COMPAT_LAST_TERM, **kw)
    xt, rt, _ = oracle.solve(ptr, col, val, b, "gmres", **kw)
    assert rc["status"] == rt["status"] == 2 and rc["steps"] == m
    assert np.linalg.norm(xc - last) <= 1e-10 * np.linalg.norm(last)
    assert np.linalg.norm(xt - full) <= 1e-10 * np.linalg.norm(full)
    # an off-by-one term or the full sum is far from it
    assert np.linalg.norm(full - last) > 1e-2 * np.linalg.norm(last)
    assert np.linalg.norm(y[m - 2] * V[:, m - 2] - last) > 1e-2 * np.linalg.norm(last)


@pytest.mark.parametrize("compat", [0, _abi.KS_COMPAT_LAST_TERM])
def test_vpgmres_preconditioner_output(oracle, compat):
    # VPGMRES's inner GMRES exits at its limit (-l L: L-1 steps) and returns
    # z = M(v_0) (vpgmres.cpp:840): with Scalar_ax's assignment only the last
    # Krylov term y_{m-1} v_{m-1}, else V_m y.  After one outer step x =
    # y_0 z with y_0 = ||b|| h00 / (h00^2 + h10^2), the one-column least
    # squares of w = A z against v_0 = b / ||b||.
    ptr, col, val, b, D = _tiny_nonsym()
    m = 5
    v0 = b / np.linalg.norm(b)
    V, y = _gmres_parts(D, v0, m)
    z = y[m - 1] * V[:, m - 1] if compat else V[:, :m] @ y
    w = D @ z
    h00 = w @ v0
    h10 = np.linalg.norm(w - h00 * v0)
    xr = np.linalg.norm(b) * h00 / (h00 ** 2 + h10 ** 2) * z
    x, r, _ = oracle.solve(ptr, col, val, b, "vpgmres", "gmres", outer_eps=0.0, outer_restart=20, outer_maxloop=2,
                           inner_maxloop=m + 1, inner_restart=20, inner_eps=0.0, ref_compat=compat)
    assert r["steps"] == 1 and r["inner_steps"] == m
    assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)


def test_kskipbicg_equals_bicg(oracle):
    ptr, col, val, b, _ = P.convdiff3d(8)
    _, rb, hb = oracle.solve(ptr, col, val, b, "bicg", outer_eps=1e-10)
    _, rk, hk = oracle.solve(ptr, col, val, b, "kskipbicg", outer_eps=1e-10, outer_kskip=2)
    at = dict(zip(hb[:, 0].astype(int), hb[:, 1]))
    for idx, v in hk[:5]:
        assert abs(v - at[int(idx)]) <= 1e-6 * at[int(idx)]


def test_vpgmres_scalar_inner_equals_gmres(oracle):
    # one inner CG step returns z = alpha v: the flexible basis spans the
    # same Krylov space, so the residual history is GMRES's
    ptr, col, val, b, _ = P.convdiff3d(7)
    _, rg, hg = oracle.solve(ptr, col, val, b, "gmres", outer_eps=1e-10, outer_restart=21, outer_maxloop=400)
    _, rv, hv = oracle.solve(ptr, col, val, b, "vpgmres", "cg", outer_eps=1e-10, outer_restart=21,
                             outer_maxloop=400, inner_maxloop=1, inner_eps=1e-30)
    m = min(len(hg), len(hv), 40)
    assert np.allclose(hg[:m, 1], hv[:m, 1], rtol=1e-7)


@pytest.mark.parametrize("compat", [0, _abi.KS_COMPAT_VPCG_FR])
def test_vpcg_scalar_inner_equals_cg(oracle, compat):
    ptr, col, val, b, _ = P.convdiff3d(7, pe=0.0)
    _, rc, hc = oracle.solve(ptr, col, val, b, "cg", outer_eps=1e-10)
    _, rv, hv = oracle.solve(ptr, col, val, b, "vpcg", "cg", outer_eps=1e-10, inner_maxloop=1, inner_eps=1e-30,
                             ref_compat=compat)
    m = min(len(hc), len(hv), 25)
    assert np.allclose(hc[:m, 1], hv[:m, 1], rtol=1e-7)


def test_mixed_inner_reaches_fp64_tolerance(oracle):
    ptr, col, val, b, xt = P.convdiff3d(10)
    x, r, h = oracle.solve(ptr, col, val, b, "vpgmres", "gmres", outer_eps=1e-10, outer_restart=21,
                           outer_maxloop=300, inner_maxloop=11, inner_restart=12, inner_precision=32,
                           inner_eps=0.0)
    assert r["status"] == 0 and r["relres_true"] <= 1.5e-10
    # -l 11 trips at count+1 >= 11: 10 inner steps per outer step
    assert r["inner_steps"] == 10 * r["steps"]


def test_maxloop_semantics(oracle):
    ptr, col, val, b, _ = P.poisson2d(16)
    _, r, h = oracle.solve(ptr, col, val, b, "cg", outer_eps=1e-15, outer_maxloop=5)
    assert r["status"] == 2 and r["iters"] == 6 and len(h) == 5
    _, r, h = oracle.solve(ptr, col, val, b, "kskipcg", outer_eps=1e-15, outer_maxloop=9, outer_kskip=2)
    assert r["status"] == 2 and r["iters"] == 9 - 2 + 1 + 0 * len(h)


# -------------------------------------------------------------- adjoint ----
def test_adjoint_diagonal_closed_form(oracle):
    d = np.linspace(1.0, 3.0, 50)
    ptr, col, val = _csr(sp.diags(d))
    b = np.random.default_rng(6).standard_normal(50)
    xt = np.random.default_rng(7).standard_normal(50)
    x = b / d
    g = x - xt
    gb, gv, _ = oracle.adjoint(ptr, col, val, x, g, "cg", outer_eps=1e-14)
    assert np.allclose(gb, g / d, rtol=1e-13)
    assert np.allclose(gv, -(g / d) * x, rtol=1e-13)


def test_adjoint_dense_and_fd(oracle):
    ptr, col, val, b, _ = P.convdiff3d(5)
    n = len(ptr) - 1
    D = P.dense(ptr, col, val)
    xt = np.random.default_rng(8).standard_normal(n)
    x = np.linalg.solve(D, b)
    gb, gv, _ = oracle.adjoint(ptr, col, val, x, x - xt, "gmres", outer_eps=1e-14, outer_restart=200)
    lam = np.linalg.solve(D.T, x - xt)
    assert np.allclose(gb, lam, rtol=1e-9, atol=1e-12)
    # central finite differences of L = 0.5||x - xt||^2 along value directions
    rng = np.random.default_rng(9)
    rows = np.repeat(np.arange(n), np.diff(ptr))
    for _ in range(4):
        dv = rng.standard_normal(len(val))
        eps = 1e-6
        Lp = 0.5 * np.sum((np.linalg.solve(P.dense(ptr, col, val + eps * dv), b) - xt) ** 2)
        Lm = 0.5 * np.sum((np.linalg.solve(P.dense(ptr, col, val - eps * dv), b) - xt) ** 2)
        fd = (Lp - Lm) / (2 * eps)
        ad = gv @ dv
        assert abs(fd - ad) <= 1e-6 * abs(ad) + 1e-9
    assert np.array_equal(gv, -gb[rows] * x[col])
