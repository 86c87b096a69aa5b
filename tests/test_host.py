"""Host-side checks of libksolve that need no GPU: the library loads and
exports every entry point include/ksolve.h declares, parameter defaults and
struct layouts, the synthetic-input generators, CRS text I/O and the Solver
command-line tool's error paths."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

from solver_cpp_amd import _abi, _lib, gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    hdr = open(os.path.join(ROOT, "include", "ksolve.h")).read()
    names = set(re.findall(r"^\s*(?:ks_status|void|int32_t|const char\*|void\*)\s*\*?\s*(ks_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 30
    lib = C.CDLL(_lib.LIB_PATH)
    for nm in names:
        assert hasattr(lib, nm), nm
    # the Python binding declares exactly the header's entry points
    assert names == set(_lib.EXPORTED)


def test_params_defaults_and_layout():
    lib = _lib.load()
    p = _abi.KsParams()
    lib.ks_params_default(C.byref(p))
    q = _abi.make_params("cg")
    for f, _ in _abi.KsParams._fields_:
        if f == "outer":
            continue
        assert getattr(p, f) == getattr(q, f), f
    # reference defaults (solver_collection.cpp:20-49)
    assert (p.outer_maxloop, p.outer_eps, p.outer_restart, p.outer_kskip, p.outer_fix) == (10000, 1e-8, 1000, 2, 2)
    assert (p.inner_maxloop, p.inner_eps, p.inner_restart, p.inner_kskip, p.inner_fix) == (50, 1e-1, 1000, 2, 2)
    assert p.inner == b"NO" and p.omp_threads == 8
    assert lib.ks_abi_version() == 1
    assert lib.ks_status_string(0) == b"converge" and lib.ks_status_string(2) == b"not converge"
    with pytest.raises(KeyError):
        _abi.make_params("cg", outer_maxlop=3)


def test_no_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    h = C.c_void_p()
    st = lib.ks_ctx_create(0, C.byref(h))
    assert st != 0 and not h.value


# ------------------------------------------------------------ generators --
def test_stencil_counts_and_values():
    for g in (3, 8, 17):
        ptr, col, val = gen.stencil(2, g, g)
        assert len(col) == 5 * g * g - 4 * g
        ptr3, col3, val3 = gen.stencil(3, g, g, g, pe=0.5)
        assert len(col3) == 7 * g ** 3 - 6 * g ** 2
    # sizes of the benchmark problems
    assert gen.stencil(2, 256, 256)[1].size == 326_656
    ptr, col, val = gen.stencil(3, 4, 4, 4, pe=0.5, vel=(1, 0.5, 0.25))
    c = np.array([1.0, 0.5, 0.25])
    D = {(i, col[j]): val[j] for i in range(64) for j in range(ptr[i], ptr[i + 1])}
    i = 1 + 4 * (1 + 4 * 1)  # interior point (1,1,1)
    assert D[(i, i)] == 6 + c.sum()
    assert D[(i, i - 1)] == -1 - c[0] and D[(i, i - 4)] == -1 - c[1] and D[(i, i - 16)] == -1 - c[2]
    assert D[(i, i + 1)] == D[(i, i + 4)] == D[(i, i + 16)] == -1.0
    for r in range(64):
        assert np.all(np.diff(col[ptr[r]:ptr[r + 1]]) > 0)
    # Pe = 0 is the symmetric Poisson matrix
    ptr, col, val = gen.stencil(3, 5, 4, 3)
    from tests._problems import dense
    A = dense(ptr, col, val)
    assert np.array_equal(A, A.T) and np.all(np.diag(A) == 6)


def test_aniso2d_spd_and_linear_in_kappa():
    from tests._problems import dense
    ptr, col, val, kap = gen.aniso2d(6, 5, 4)
    A = dense(ptr, col, val)
    assert np.array_equal(A, A.T)
    assert np.linalg.eigvalsh(A).min() > 0
    assert len(kap) == 7 * 5 + 6 * 6
    assert np.allclose(gen.aniso2d_values(6, 5, kap), val)
    d = np.random.default_rng(0).standard_normal(len(kap))
    assert np.allclose(gen.aniso2d_values(6, 5, kap + 0.5 * d), val + 0.5 * gen.aniso2d_values(6, 5, d))
    # benchmark size: n = 1024^2, nnz = 5,238,784
    n = 1024
    assert n * n + 2 * ((n - 1) * n + n * (n - 1)) == 5_238_784


def test_zipf_matrix_properties():
    ptr, col, val = gen.zipf(200_000, 1.5, 4096, (51, 52, 53))
    L = np.diff(ptr)
    assert L.min() >= 1 and L.max() <= 4096
    # truncated Zipf(1.5) on 1..4096: mean sum k^-0.5 / sum k^-1.5 = 49.0
    k = np.arange(1, 4097, dtype=np.float64)
    mean = (k ** -0.5).sum() / (k ** -1.5).sum()
    assert abs(mean - 49.0) < 0.05
    assert abs(L.mean() - mean) < 3.0
    rows = np.repeat(np.arange(200_000), L)
    for r in range(0, 200_000, 997):
        c = col[ptr[r]:ptr[r + 1]]
        assert np.all(np.diff(c) > 0) and r in c
    diag = val[col == rows]
    off = np.abs(val).copy()
    off[col == rows] = 0
    rs = np.bincount(rows, weights=off, minlength=200_000)
    assert np.allclose(diag, np.where(L > 1, 1.1 * rs, 1.0))
    # deterministic
    ptr2, col2, val2 = gen.zipf(200_000, 1.5, 4096, (51, 52, 53))
    assert np.array_equal(col, col2) and np.array_equal(val, val2)


# Independent restatement of the published generators: SplitMix64 (Steele,
# Lea, Flood, OOPSLA 2014) and xoshiro256** 1.0 (Blackman, Vigna 2018).
M64 = (1 << 64) - 1


def _splitmix(s):
    while True:
        s = (s + 0x9E3779B97F4A7C15) & M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        yield z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M64


def _xoshiro(seed):
    sm = _splitmix(seed)
    s = [next(sm) for _ in range(4)]
    while True:
        r = (_rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = _rotl(s[3], 45)
        yield r


def test_rng_matches_published_algorithms():
    # SplitMix64 from state 0: first output 0xE220A8397B1DCDAF
    assert next(_splitmix(0)) == 0xE220A8397B1DCDAF
    for seed in (1, 3, 41):
        g = _xoshiro(seed)
        u = np.array([(next(g) >> 11) * 2.0 ** -53 for _ in range(64)])
        assert np.array_equal(gen.vector(seed, 64, "uniform", 0.0, 1.0), u)
        assert np.array_equal(gen.vector(seed, 64, "uniform", -1.0, 1.0), -1.0 + 2.0 * u)
    # Box-Muller normals, both outputs in turn
    g = _xoshiro(2)
    z = []
    for _ in range(16):
        u1 = 1.0 - (next(g) >> 11) * 2.0 ** -53
        u2 = (next(g) >> 11) * 2.0 ** -53
        r = math.sqrt(-2.0 * math.log(u1))
        z += [r * math.cos(2 * math.pi * u2), r * math.sin(2 * math.pi * u2)]
    assert np.allclose(gen.vector(2, 32, "normal"), z, rtol=1e-15, atol=1e-15)
    big = gen.vector(2, 400_000, "normal")
    assert abs(big.mean()) < 0.01 and abs(big.std() - 1) < 0.01


# ------------------------------------------------------------- CRS text --
def test_crs_roundtrip_and_header_checks(tmp_path):
    ptr, col, val = gen.stencil(3, 5, 4, 3, pe=0.8)
    n = len(ptr) - 1
    b = gen.vector(3, n, "uniform")
    x = gen.vector(4, n, "normal")
    d = str(tmp_path / "M")
    gen.write_crs(d, ptr, col, val, b, x)
    p2, c2, v2, b2, x2 = gen.read_crs(d)
    assert np.array_equal(p2, ptr) and np.array_equal(c2, col) and np.array_equal(v2, val)
    assert np.array_equal(b2, b) and np.array_equal(x2, x)
    first = open(os.path.join(d, "ColVal.txt")).readline().split()
    assert first == [str(n), str(n), str(len(col))]
    # the three headers must agree (solver_collection.cpp:701-719)
    txt = open(os.path.join(d, "Ptr.txt")).read().split("\n", 1)
    open(os.path.join(d, "Ptr.txt"), "w").write(f"{n} {n} {len(col) + 1}\n" + txt[1])
    with pytest.raises(RuntimeError):
        gen.read_crs(d)
    os.remove(os.path.join(d, "bx.txt"))
    with pytest.raises(RuntimeError):
        gen.read_crs(d)


def _write_messy_crs(d, ptr, col, val, b, x):
    # the reference's fscanf accepts any whitespace: tabs, runs of blanks,
    # CRLF line ends and several records on one line
    os.makedirs(d, exist_ok=True)
    n, nnz = len(ptr) - 1, len(col)
    hdr = f"{n}  {n}\t{nnz}\r\n"
    recs = [f"{c}\t {float(v)!r}" for c, v in zip(col, val)]
    body = "".join(r + ("  " if i % 3 else "\r\n") for i, r in enumerate(recs))
    open(os.path.join(d, "ColVal.txt"), "w", newline="").write(hdr + body + "\n")
    open(os.path.join(d, "Ptr.txt"), "w", newline="").write(hdr + " ".join(str(int(p)) for p in ptr) + "\n")
    bx = "\n".join(f"{float(bi)!r}   {float(xi)!r}" for bi, xi in zip(b, x))
    open(os.path.join(d, "bx.txt"), "w", newline="").write(hdr + bx + "\n\n")


def test_crs_parallel_parse_any_whitespace(tmp_path):
    # > 1 MiB of text, so the parser splits it across threads
    ptr, col, val = gen.stencil(3, 24, 24, 24, pe=0.5)
    n = len(ptr) - 1
    b = gen.vector(7, n, "uniform")
    x = gen.vector(8, n, "normal")
    d = str(tmp_path / "messy")
    _write_messy_crs(d, ptr, col, val, b, x)
    assert os.path.getsize(os.path.join(d, "ColVal.txt")) > (1 << 20)
    p2, c2, v2, b2, x2 = gen.read_crs(d)
    # repr() round-trips exactly, so the parse must be bit-exact
    assert np.array_equal(p2, ptr) and np.array_equal(c2, col) and np.array_equal(v2, val)
    assert np.array_equal(b2, b) and np.array_equal(x2, x)
    # a malformed token is an error, not a silent truncation
    hdr, body = open(os.path.join(d, "Ptr.txt")).read().split("\n", 1)
    toks = body.split(" ")
    toks[5] += "x"
    open(os.path.join(d, "Ptr.txt"), "w").write(hdr + "\n" + " ".join(toks))
    with pytest.raises(RuntimeError):
        gen.read_crs(d)


@pytest.mark.parametrize("which,token", [("col", "4294967301"), ("col", "-1"), ("ptr", "2147483648"),
                                         ("col", "99999999999999999999999")])
def test_crs_index_out_of_int32_range(tmp_path, which, token):
    # an index that does not fit the int32 CSR arrays is a malformed file:
    # 4294967301 must not wrap to 5 and pass the column-range check as a
    # different valid matrix
    ptr, col, val = gen.stencil(2, 6, 5)
    n = len(ptr) - 1
    d = str(tmp_path / "M")
    gen.write_crs(d, ptr, col, val, gen.vector(3, n, "uniform"))
    fn = os.path.join(d, "ColVal.txt" if which == "col" else "Ptr.txt")
    lines = open(fn).read().split("\n")
    f = lines[3].split()
    f[0] = token
    lines[3] = " ".join(f)
    open(fn, "w").write("\n".join(lines))
    with pytest.raises(RuntimeError):
        gen.read_crs(d)


def test_crs_binary_sidecar(tmp_path):
    ptr, col, val = gen.stencil(3, 9, 8, 7, pe=0.8)
    n = len(ptr) - 1
    b = gen.vector(3, n, "uniform")
    x = gen.vector(4, n, "normal")
    d = str(tmp_path / "M")
    gen.write_crs(d, ptr, col, val, b, x)
    st = gen.crs_stamp(d)
    f = os.path.join(d, "CRS.ksb")
    gen.write_crs_bin(f, ptr, col, val, b, x, stamp=st)
    p2, c2, v2, b2, x2, st2 = gen.read_crs_bin(f)
    assert st2 == st
    for u, v in ((p2, ptr), (c2, col), (v2, val), (b2, b), (x2, x)):
        assert np.array_equal(u, v)
    # the stamp follows the text files
    txt = open(os.path.join(d, "bx.txt")).read()
    open(os.path.join(d, "bx.txt"), "w").write(txt + "\n")
    assert gen.crs_stamp(d) != st
    # one flipped payload byte, a truncated file, a foreign header: all rejected
    raw = bytearray(open(f, "rb").read())
    for bad in (raw[:64] + raw[64:80] + bytes([raw[80] ^ 1]) + raw[81:], raw[:-8], b"NOTKSCSR" + raw[8:]):
        open(f, "wb").write(bytes(bad))
        with pytest.raises(RuntimeError):
            gen.read_crs_bin(f)
    # without b / x: read as zeros
    gen.write_crs_bin(f, ptr, col, val)
    p3, c3, v3, b3, x3, st3 = gen.read_crs_bin(f)
    assert st3 == 0 and np.array_equal(v3, val) and not b3.any() and not x3.any()


# --------------------------------------------------------------- the CLI --
SOLVER = os.path.join(ROOT, "solver_cpp_amd", "Solver")


def _cli(*args, cwd=None):
    return subprocess.run([SOLVER, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


def test_cli_flag_errors(tmp_path):
    assert os.path.exists(SOLVER)
    r = _cli("-m", "x")
    assert r.returncode != 0 and "OuterSolver" in r.stderr
    r = _cli("-S", "nosuch", "-m", "x")
    assert r.returncode != 0 and "OuterSolverName can not found in List" in r.stderr
    r = _cli("-S", "cg")
    assert r.returncode != 0 and "Must set one input source" in r.stderr
    r = _cli("-S", "cg", "-m", "a", "-M", "b")
    assert r.returncode != 0 and "Only can set one input source" in r.stderr
    r = _cli("-S", "cg", "-m", "missing", "-p", str(tmp_path))
    assert r.returncode != 0 and "CRS" in r.stderr
    r = _cli("-S", "cg", "-m", "x", "-g")
    assert r.returncode != 0 and "one GPU" in r.stderr
    r = _cli("--bogus")
    assert r.returncode != 0
    r = _cli("-h")
    assert r.returncode == 0 and "OuterSolver" in r.stderr


def test_cli_header_block_as_reference(tmp_path):
    # collection::showCMD's block (solver_collection.cpp:603-651) is printed
    # before the GPU is opened; the reference's log scraper reads its
    # "CUDA :" and "MultiGPU :" lines (data2csv.sh:6-15)
    from tests._scrape import seek
    ptr, col, val = gen.stencil(2, 6, 6, 1, pe=0.0)
    root = tmp_path / "root"
    gen.write_crs(str(root / "Matrix" / "CRS" / "P6"), ptr, col, val, gen.vector(1, 36, "uniform"))
    work = tmp_path / "work"
    (work / "output").mkdir(parents=True)
    r = _cli("-S", "vpgmres", "-s", "gmres", "-R", "21", "-l", "11", "-m", "P6", "-p", str(root), cwd=str(work))
    lines = r.stdout.splitlines()
    for want in ("N : 36", "NNZ : 156", "VP : On", "CUDA : On", "Pinned : Off", "MultiGPU : Off",
                 "OuterSolver : VPGMRES", "OuterRestart : 21", "InnerSolver : GMRES", "InnerMaxLoop : 11",
                 "OpenMP thread : 8"):
        assert want in lines, want
    f = seek(r.stdout)
    assert f["is_gpu"] and not f["is_multigpu"]
    # nothing in the block may be taken for the fields scraped from the solve
    assert f["Total"] is None and f["Loop"] is None and f["CPU-Dot"] is None and f["Other"] is None


def test_cli_bin_cache_written_then_used(tmp_path):
    # the matrix is read (and the cache written) before the GPU is opened, so
    # this runs without a GPU: the solve itself then fails with "no GPU"
    ptr, col, val = gen.stencil(2, 12, 12, 1, pe=0.0)
    n = len(ptr) - 1
    root = tmp_path / "root"
    d = root / "Matrix" / "CRS" / "P12"
    gen.write_crs(str(d), ptr, col, val, gen.vector(1, n, "uniform"))
    work = tmp_path / "work"
    (work / "output").mkdir(parents=True)
    args = ("-S", "cg", "-m", "P12", "-p", str(root), "--bin-cache")
    r1 = _cli(*args, cwd=str(work))
    assert os.path.exists(d / "CRS.ksb")
    assert "InputSource : CRS\n" in r1.stdout
    r2 = _cli(*args, cwd=str(work))
    assert "InputSource : CRS (binary cache)" in r2.stdout
    p2, c2, v2, *_ = gen.read_crs_bin(str(d / "CRS.ksb"))
    assert np.array_equal(p2, ptr) and np.array_equal(c2, col) and np.array_equal(v2, val)
