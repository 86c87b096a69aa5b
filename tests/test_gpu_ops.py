"""GPU parity of the BLAS-layer kernels against the CPU oracle, through the C ABI.

Tolerances: a row sum of m terms differs from the oracle's sequential sum by
at most ~m*eps*sum|a_ij x_j| (different order, FMA), so SpMV is checked
against 4*m*eps*sum|a x|; reductions against 64*eps*sum|x_i y_i|; the
transpose (pure integer/permutation work) bit for bit.
"""
import numpy as np
import pytest

from tests import _problems as P

pytestmark = pytest.mark.gpu

EPS64 = 2.0 ** -53
EPS32 = 2.0 ** -24


def _mat(ctx, ptr, col, val, **kw):
    from solver_cpp_amd import api
    return api.Matrix(ctx, ptr, col, val, **kw)


LAYOUTS = ["auto", "col32", "delta", "stream", "wg", "wg_col32", "wg_delta", "wave", "vector", "merge", "panels"]


def _auto_kind(ptr):
    # the library's default: merge path for heavy-tailed row lengths
    rl = np.diff(ptr)
    return 2 if rl.max() > 256 and rl.max() > 16 * rl.mean() else 0


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("case", ["poisson2d", "convdiff3d", "zipf"])
def test_spmv_fp64_fp32_and_transpose(ctx, oracle, case, layout):
    if case == "poisson2d":
        ptr, col, val, _, _ = P.poisson2d(40)
    elif case == "convdiff3d":
        ptr, col, val, _, _ = P.convdiff3d(13)
    else:
        ptr, col, val, _ = P.zipf_small(5000, 600)
    n = len(ptr) - 1
    x = np.random.default_rng(7).standard_normal(n)
    A = _mat(ctx, ptr, col, val, transpose=True, fp32=True, layout=layout)
    kinds = {"auto": _auto_kind(ptr), "col32": _auto_kind(ptr), "delta": _auto_kind(ptr), "stream": 0, "wg": 0,
             "wg_col32": 0, "wg_delta": 0, "wave": 0, "vector": 1, "merge": 2, "panels": 3}
    assert A.info()["spmv_kind"] == kinds[layout]
    m = np.diff(ptr).max()
    y = A.spmv(x)
    yo = oracle.spmv(ptr, col, val, x)
    bound = 4 * m * EPS64 * P.row_abs_bound(ptr, col, val, x) + 1e-300
    assert np.all(np.abs(y - yo) <= bound)
    # transpose: bit-exact against a stable counting sort
    tp, tc, tv = A.transpose_csr()
    otp, otc, otv = oracle.transpose(ptr, col, val)
    assert np.array_equal(tp, otp) and np.array_equal(tc, otc) and np.array_equal(tv, otv)
    yt = A.spmv(x, trans=True)
    yto = oracle.spmv(otp, otc, otv, x)
    assert np.all(np.abs(yt - yto) <= 4 * np.diff(otp).max() * EPS64 * P.row_abs_bound(otp, otc, otv, x) + 1e-300)
    # adjoint identity y^T (A x) = (A^T y)^T x
    z = np.random.default_rng(8).standard_normal(n)
    lhs, rhs = z @ A.spmv(x), A.spmv(z, trans=True) @ x
    assert abs(lhs - rhs) <= 1e-12 * (np.abs(z) @ P.row_abs_bound(ptr, col, val, x))
    # FP32 values and vectors
    x32 = x.astype(np.float32)
    y32 = A.spmv(x32, precision=32)
    yo32 = oracle.spmv_f32(ptr, col, val.astype(np.float32), x32)
    b32 = 4 * m * EPS32 * P.row_abs_bound(ptr, col, val.astype(np.float32).astype(np.float64), x32) + 1e-30
    assert np.all(np.abs(y32.astype(np.float64) - yo32.astype(np.float64)) <= b32)


@pytest.mark.parametrize("layout", ["stream", "wg", "wg_col32", "wg_delta", "wave"])
@pytest.mark.parametrize("case", ["poisson2d", "convdiff3d", "aniso"])
def test_spmv_bitexact_short_rows(ctx, oracle, case, layout):
    # rows of <= 8 entries are summed sequentially in CSR order from products
    # rounded in the value type: the oracle's (the reference's) exact order
    if case == "poisson2d":
        ptr, col, val, _, _ = P.poisson2d(41)
    elif case == "convdiff3d":
        ptr, col, val, _, _ = P.convdiff3d(13)
    else:
        ptr, col, val, _ = P.aniso(37)
    n = len(ptr) - 1
    x = np.random.default_rng(11).standard_normal(n)
    A = _mat(ctx, ptr, col, val, fp32=True, layout=layout)
    assert np.array_equal(A.spmv(x), oracle.spmv(ptr, col, val, x))
    x32 = x.astype(np.float32)
    assert np.array_equal(A.spmv(x32, precision=32), oracle.spmv_f32(ptr, col, val.astype(np.float32), x32))


def _band_csr(n, offsets, seed):
    # row i holds the entries i + d, d in offsets, that fall inside [0, n)
    rows = [[i + d for d in sorted(offsets) if 0 <= i + d < n] for i in range(n)]
    ptr = np.zeros(n + 1, np.int32)
    ptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in r], np.int32)
    val = np.random.default_rng(seed).standard_normal(len(col))
    return ptr, col, val


@pytest.mark.parametrize("reach", [32767, 32768])
def test_spmv_col16_offset_limit(ctx, oracle, reach):
    # entries at row -+ reach: within int16 offsets at 32767 (the SpMV then
    # reads col - row as int16, A^T too), delta-coded columns at 32768 (and
    # int32 columns when those are not allowed either); all bit-exact
    # against the oracle (rows of <= 8 entries) and equal to the int32 path
    n = 70_000
    ptr, col, val = _band_csr(n, [-reach, -3, 0, 1, reach], 21)
    A = _mat(ctx, ptr, col, val, transpose=True, fp32=True)
    B = _mat(ctx, ptr, col, val, transpose=True, fp32=True, layout="col32")
    ib = 2
    for prec in (64, 32):
        assert A.index_bytes(False, prec) == ib and A.index_bytes(True, prec) == ib
        assert B.index_bytes(False, prec) == 4 and B.index_bytes(True, prec) == 4
    x = np.random.default_rng(22).standard_normal(n)
    y = A.spmv(x)
    assert np.array_equal(y, oracle.spmv(ptr, col, val, x)) and np.array_equal(y, B.spmv(x))
    otp, otc, otv = oracle.transpose(ptr, col, val)
    yt = A.spmv(x, trans=True)
    assert np.array_equal(yt, oracle.spmv(otp, otc, otv, x)) and np.array_equal(yt, B.spmv(x, trans=True))
    x32 = x.astype(np.float32)
    y32 = A.spmv(x32, precision=32)
    assert np.array_equal(y32, oracle.spmv_f32(ptr, col, val.astype(np.float32), x32))
    assert np.array_equal(y32, B.spmv(x32, precision=32))


def test_spmv_col16_long_rows(ctx, oracle):
    # rows longer than a workgroup tile (summed in place, not staged) read
    # with int16 offsets: same sums as the int32 path
    ptr, col, val, _ = P.zipf_small(20_000, 4096)
    assert np.diff(ptr).max() > 2048
    A = _mat(ctx, ptr, col, val, fp32=True, layout="wg")
    B = _mat(ctx, ptr, col, val, fp32=True, layout="wg_col32")
    # delta-coded columns do not apply (tiles of one long row): int16 offsets
    C = _mat(ctx, ptr, col, val, fp32=True, layout="wg_delta")
    assert A.index_bytes() == 2 and B.index_bytes() == 4 and C.index_bytes() == 2
    x = np.random.default_rng(23).standard_normal(len(ptr) - 1)
    assert np.array_equal(A.spmv(x), B.spmv(x)) and np.array_equal(C.spmv(x), B.spmv(x))
    x32 = x.astype(np.float32)
    assert np.array_equal(A.spmv(x32, precision=32), B.spmv(x32, precision=32))
    assert np.array_equal(C.spmv(x32, precision=32), B.spmv(x32, precision=32))


def _check_same_as_int32(ctx, oracle, ptr, col, val, A, seed):
    # A's SpMV (A and A^T, FP64 and FP32) bit-exact against the oracle (rows
    # of <= 8 entries: the oracle's order) and equal to the int32 columns'
    B = _mat(ctx, ptr, col, val, transpose=True, fp32=True, layout="col32")
    n = len(ptr) - 1
    x = np.random.default_rng(seed).standard_normal(n)
    y = A.spmv(x)
    assert np.array_equal(y, oracle.spmv(ptr, col, val, x)) and np.array_equal(y, B.spmv(x))
    otp, otc, otv = oracle.transpose(ptr, col, val)
    yt = A.spmv(x, trans=True)
    assert np.array_equal(yt, oracle.spmv(otp, otc, otv, x)) and np.array_equal(yt, B.spmv(x, trans=True))
    x32 = x.astype(np.float32)
    y32 = A.spmv(x32, precision=32)
    assert np.array_equal(y32, oracle.spmv_f32(ptr, col, val.astype(np.float32), x32))
    assert np.array_equal(y32, B.spmv(x32, precision=32))
    yt32 = A.spmv(x32, precision=32, trans=True)
    assert np.array_equal(yt32, oracle.spmv_f32(otp, otc, otv.astype(np.float32), x32))
    assert np.array_equal(yt32, B.spmv(x32, precision=32, trans=True))
    B.close()


@pytest.mark.parametrize("offsets,ib,ibt", [
    ([-49152, -3, 0, 1, 49152], 2, 2),   # first column at row - 49152: the lowest code
    ([-49153, -3, 0, 1, 49153], 4, 4),   # one below
    ([0, 65535], 2, 4),                  # gap 65535 in A; A^T's first column at row - 65535
    ([0, 65536], 4, 4),                  # gap 65536
    ([16383, 40000], 2, 2),              # first column at row + 16383: the highest code
    ([16384, 40000], 4, 2),              # one above in A; A^T's columns at row - 40000, row - 16384
])
def test_spmv_delta_columns_limits(ctx, oracle, offsets, ib, ibt):
    # offsets beyond int16: the SpMV reads delta-coded 16-bit columns where
    # every row encodes (the first column within [row - 49152, row + 16383],
    # sorted, gaps below 65536), A and A^T separately, else int32 columns;
    # the same sums either way
    n = 140_000
    ptr, col, val = _band_csr(n, offsets, 24)
    A = _mat(ctx, ptr, col, val, transpose=True, fp32=True)
    for prec in (64, 32):
        assert A.index_bytes(False, prec) == ib and A.index_bytes(True, prec) == ibt
    _check_same_as_int32(ctx, oracle, ptr, col, val, A, 25)
    A.close()


def test_spmv_delta_columns_unsorted_rows(ctx, oracle):
    # rows of A in descending column order (a negative gap) keep int32
    # columns; A^T, built sorted, still gets delta-coded columns
    n = 100_000
    ptr, col, val = _band_csr(n, [-40000, -1, 0, 1, 40000], 26)
    for i in range(n):
        col[ptr[i]:ptr[i + 1]] = col[ptr[i]:ptr[i + 1]][::-1].copy()
        val[ptr[i]:ptr[i + 1]] = val[ptr[i]:ptr[i + 1]][::-1].copy()
    A = _mat(ctx, ptr, col, val, transpose=True, fp32=True)
    assert A.index_bytes(False, 64) == 4 and A.index_bytes(True, 64) == 2
    _check_same_as_int32(ctx, oracle, ptr, col, val, A, 27)
    A.close()


def test_spmv_delta_columns_tile_condition(ctx, oracle):
    # 18 entries per row: the SpMV sums a tile's rows with two lanes each,
    # where delta-coded columns (one lane per row) do not apply, so offsets
    # beyond int16 keep int32 columns, and offsets within int16 keep int16
    # offsets even when delta-coded columns are preferred
    n = 100_000
    wide = [-40000] + list(range(-8, 8)) + [40000]
    ptr, col, val = _band_csr(n, wide, 28)
    A = _mat(ctx, ptr, col, val, transpose=True, fp32=True)
    assert A.index_bytes(False, 64) == 4 and A.index_bytes(True, 64) == 4
    x = np.random.default_rng(29).standard_normal(n)
    B = _mat(ctx, ptr, col, val, transpose=True, fp32=True, layout="col32")
    assert np.array_equal(A.spmv(x), B.spmv(x)) and np.array_equal(A.spmv(x, trans=True), B.spmv(x, trans=True))
    ptr, col, val = _band_csr(n, [-30000] + list(range(-8, 8)) + [30000], 30)
    C = _mat(ctx, ptr, col, val, transpose=True, fp32=True, layout="delta")
    D = _mat(ctx, ptr, col, val, transpose=True, fp32=True, layout="col32")
    assert C.index_bytes(False, 64) == 2 and C.index_bytes(True, 64) == 2
    assert np.array_equal(C.spmv(x), D.spmv(x)) and np.array_equal(C.spmv(x, trans=True), D.spmv(x, trans=True))
    for M in (A, B, C, D):
        M.close()


@pytest.mark.parametrize("layout", LAYOUTS)
def test_spmv_edge_cases(ctx, oracle, layout):
    # empty rows, a dense row, a single entry
    n = 70
    rows = [[], [0], list(range(n)), [5, 69], []] + [[i] for i in range(5, n)]
    ptr = np.zeros(n + 1, np.int32)
    cols = []
    for i, r in enumerate(rows[:n]):
        cols += r
        ptr[i + 1] = len(cols)
    col = np.array(cols, np.int32)
    val = np.random.default_rng(3).standard_normal(len(col))
    A = _mat(ctx, ptr, col, val, transpose=True, layout=layout)
    x = np.random.default_rng(4).standard_normal(n)
    y = A.spmv(x)
    yo = oracle.spmv(ptr, col, val, x)
    assert np.allclose(y, yo, rtol=1e-13, atol=1e-13)
    assert y[0] == 0.0 and y[4] == 0.0
    tp, tc, tv = A.transpose_csr()
    otp, otc, otv = oracle.transpose(ptr, col, val)
    assert np.array_equal(tp, otp) and np.array_equal(tc, otc) and np.array_equal(tv, otv)


@pytest.mark.parametrize("layout", LAYOUTS)
def test_spmv_long_rows(ctx, oracle, layout):
    # Zipf rows up to 4096 entries: rows longer than a workgroup tile (2048),
    # a wavefront tile (512) and a merge-path segment (512 items)
    ptr, col, val, _ = P.zipf_small(20_000, 4096)
    assert np.diff(ptr).max() > 2048
    n = len(ptr) - 1
    x = np.random.default_rng(9).standard_normal(n)
    A = _mat(ctx, ptr, col, val, transpose=True, fp32=True, layout=layout)
    m = np.diff(ptr).max()
    y = A.spmv(x)
    yo = oracle.spmv(ptr, col, val, x)
    assert np.all(np.abs(y - yo) <= 4 * m * EPS64 * P.row_abs_bound(ptr, col, val, x) + 1e-300)
    yt = A.spmv(x, trans=True)
    otp, otc, otv = oracle.transpose(ptr, col, val)
    yto = oracle.spmv(otp, otc, otv, x)
    assert np.all(np.abs(yt - yto) <= 4 * np.diff(otp).max() * EPS64 * P.row_abs_bound(otp, otc, otv, x) + 1e-300)
    x32 = x.astype(np.float32)
    v32 = val.astype(np.float32)
    y32 = A.spmv(x32, precision=32)
    yo32 = oracle.spmv_f32(ptr, col, v32, x32)
    b32 = 4 * m * EPS32 * P.row_abs_bound(ptr, col, v32.astype(np.float64), x32) + 1e-30
    assert np.all(np.abs(y32.astype(np.float64) - yo32.astype(np.float64)) <= b32)


def _dense_column_csr(n, seed):
    # a diagonal plus three dense columns of A (rows of A^T longer than one
    # LDS sort of 8192 entries: 1, 2 and 4 merge passes), unsorted column
    # order inside rows and repeated (row, column) entries
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        r = [i, 0]
        if i % 5 == 0:
            r.append(1)
        if i % 12 == 0:
            r.append(2)
        if i % 997 == 3:
            r += [3, 3]
        rng.shuffle(r)
        rows.append(r)
    ptr = np.zeros(n + 1, np.int32)
    ptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.fromiter((c for r in rows for c in r), np.int32, count=int(ptr[-1]))
    val = rng.standard_normal(len(col))
    return ptr, col, val


@pytest.mark.parametrize("layout", ["auto", "stream", "panels"])
def test_transpose_dense_columns(ctx, oracle, layout):
    ptr, col, val = _dense_column_csr(100_000, 17)
    n = len(ptr) - 1
    A = _mat(ctx, ptr, col, val, transpose=True, fp32=True, layout=layout)
    tp, tc, tv = A.transpose_csr()
    otp, otc, otv = oracle.transpose(ptr, col, val)
    assert np.diff(otp)[:3].tolist() == [n + 1, 20_001, 8_335]  # the diagonal adds one each
    assert np.array_equal(tp, otp) and np.array_equal(tc, otc) and np.array_equal(tv, otv)
    x = np.random.default_rng(5).standard_normal(n)
    yt = A.spmv(x, trans=True)
    yto = oracle.spmv(otp, otc, otv, x)
    assert np.all(np.abs(yt - yto) <= 4 * np.diff(otp).max() * EPS64 * P.row_abs_bound(otp, otc, otv, x) + 1e-300)
    # refreshed values follow the same permutation
    A.update_values(3.0 * val)
    assert np.array_equal(A.transpose_csr()[2], 3.0 * otv)
    A.close()


def test_transpose_segment_sizes(ctx, oracle):
    # A^T rows of every length 0..140 (each sort path and the limits between
    # them: 16/17 and 64/65 entries), rows drawn with repeats
    rng = np.random.default_rng(23)
    n = 4000
    lens = np.arange(n) % 141
    rows = [[] for _ in range(n)]
    for j, L in enumerate(lens):
        for i in rng.integers(0, n, L):
            rows[i].append(j)
    for r in rows:
        rng.shuffle(r)
    ptr = np.zeros(n + 1, np.int32)
    ptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.fromiter((c for r in rows for c in r), np.int32, count=int(ptr[-1]))
    val = rng.standard_normal(len(col))
    A = _mat(ctx, ptr, col, val, transpose=True, layout="stream")
    tp, tc, tv = A.transpose_csr()
    otp, otc, otv = oracle.transpose(ptr, col, val)
    assert np.array_equal(np.diff(otp), lens)
    assert np.array_equal(tp, otp) and np.array_equal(tc, otc) and np.array_equal(tv, otv)
    A.close()


def test_borrowed_matrix(ctx, oracle):
    # KS_MAT_BORROW: the caller's device CSR is used in place
    import torch
    from solver_cpp_amd import api
    ptr, col, val, _, _ = P.convdiff3d(12)
    n = len(ptr) - 1
    tp, tc, tv = (torch.from_numpy(a).cuda() for a in (ptr, col, val))
    A = _mat(ctx, tp, tc, tv, transpose=True, fp32=True, borrow=True)
    x = np.random.default_rng(3).standard_normal(n)
    assert np.array_equal(A.spmv(x), oracle.spmv(ptr, col, val, x))
    # values changed in place, then made visible to the derived copies
    tv.mul_(2.0)
    A.update_values(tv)
    v2 = 2.0 * val
    assert np.array_equal(A.spmv(x), oracle.spmv(ptr, col, v2, x))
    x32 = x.astype(np.float32)
    assert np.array_equal(A.spmv(x32, precision=32), oracle.spmv_f32(ptr, col, v2.astype(np.float32), x32))
    otp, otc, otv = oracle.transpose(ptr, col, v2)
    assert np.array_equal(A.spmv(x, trans=True), oracle.spmv(otp, otc, otv, x))
    A.close()
    with pytest.raises(api.KsolveError):
        api.Matrix(ctx, ptr, col, val, borrow=True)  # host arrays cannot be borrowed


def test_bad_csr_rejected(ctx):
    from solver_cpp_amd import api
    ptr = np.array([0, 1, 2], np.int32)
    with pytest.raises(api.KsolveError):
        api.Matrix(ctx, ptr, np.array([0, 5], np.int32), np.ones(2))
    with pytest.raises(api.KsolveError):
        api.Matrix(ctx, np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32), np.ones(2))


@pytest.mark.parametrize("n", [1, 63, 1000, 100_003, 3_000_001])
def test_dot_norm_axpy(ctx, n):
    import math
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    d = ctx.dot(x, y)
    exact = math.fsum(x * y)  # products exact to 1 ulp each
    assert abs(d - exact) <= 64 * EPS64 * np.abs(x * y).sum() + 1e-300
    nr = ctx.norm2(x)
    assert abs(nr - math.sqrt(math.fsum(x * x))) <= 64 * EPS64 * nr
    # closed form: sum_{i<n} i = n(n-1)/2 (exact in FP64 at these n)
    i = np.arange(n, dtype=np.float64)
    assert ctx.dot(i, np.ones(n)) == n * (n - 1) / 2
    o = ctx.axpy(0.75, x, y)
    assert np.allclose(o, 0.75 * x + y, rtol=0, atol=4 * EPS64 * (np.abs(0.75 * x) + np.abs(y)).max())
    # FP32 inputs are accumulated in FP64
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    d32 = ctx.dot(x32, y32, precision=32)
    assert abs(d32 - math.fsum(x32.astype(np.float64) * y32)) <= 64 * EPS64 * np.abs(x32 * y32.astype(np.float64)).sum() + 1e-300


def test_deterministic_reductions(ctx):
    x = np.random.default_rng(1).standard_normal(2_000_000)
    vals = {ctx.dot(x, x) for _ in range(5)}
    assert len(vals) == 1


@pytest.mark.parametrize("n", [50_000, 50_001])  # 16-byte vector path / scalar path
@pytest.mark.parametrize("k", [0, 3, 20, 40])
def test_cgs_step(ctx, oracle, k, n):
    rng = np.random.default_rng(k)
    V = np.linalg.qr(rng.standard_normal((n, k + 1)))[0].T.copy()
    w = rng.standard_normal(n)
    wg, hg = ctx.cgs(V, w)
    wo, ho = oracle.cgs(V, w, k)
    assert np.allclose(hg, ho, rtol=1e-12, atol=1e-12 * np.linalg.norm(w))
    assert np.allclose(wg, wo, rtol=0, atol=1e-12 * np.linalg.norm(w))
    # invariant: w is orthogonal to the basis afterwards
    assert np.abs(V @ wg).max() <= 1e-12 * np.linalg.norm(w)


@pytest.mark.parametrize("layout", ["stream", "wg", "wave", "merge", "panels"])
@pytest.mark.parametrize("k", [1, 2, 4])
def test_kskip_basis_and_gram(ctx, oracle, k, layout):
    ptr, col, val, _, _ = P.convdiff3d(10, pe=0.0)
    n = len(ptr) - 1
    rng = np.random.default_rng(k)
    r, p = rng.standard_normal(n), rng.standard_normal(n)
    A = _mat(ctx, ptr, col, val, layout=layout)
    Ar, Ap = A.kskip_basis(k, r, p)
    Aro, Apo = oracle.kskip_basis(ptr, col, val, k, r, p)
    for G, O in ((Ar, Aro), (Ap, Apo)):
        for i in range(G.shape[0]):
            assert np.linalg.norm(G[i] - O[i]) <= 1e-13 * np.linalg.norm(O[i])
    d, e, z = ctx.kskip_gram(k, r, p, Aro, Apo)
    do, eo, zo = oracle.kskip_gram(k, r, p, Aro, Apo)
    sc = np.linalg.norm(r) * np.linalg.norm(Apo, axis=1).max() + np.linalg.norm(p) * np.linalg.norm(Apo, axis=1).max()
    for g, o in ((d, do), (e, eo), (z, zo)):
        assert np.allclose(g, o, rtol=0, atol=1e-13 * sc)


def test_sddmm(ctx, oracle):
    ptr, col, val, _, _ = P.convdiff3d(9)
    n = len(ptr) - 1
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    A = _mat(ctx, ptr, col, val)
    g = A.sddmm(-1.0, u, v)
    o = oracle.sddmm(ptr, col, -1.0, u, v)
    assert np.array_equal(g, o)


def test_bench_stream(ctx):
    # the streaming reference kernels run and report a rate; sizes that are
    # not whole 16-byte chunks are refused
    from solver_cpp_amd import api
    for kind in ("read", "copy"):
        assert ctx.stream_rate(kind, nbytes=64 << 20, reps=2) > 0
    with pytest.raises(api.KsolveError):
        ctx.stream_rate("read", nbytes=(64 << 20) + 8)
