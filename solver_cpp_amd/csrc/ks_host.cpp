// ks_host.cpp - host-only entry points of libksolve: parameter defaults,
// status strings, deterministic synthetic inputs, and the reference's CRS
// text files (solver_collection.cpp:655-828).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>

#include "../../include/ksolve.h"

namespace {

// SplitMix64 (Steele, Lea, Flood 2014) - seeds xoshiro256**.
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
};

// xoshiro256** (Blackman, Vigna 2018).
struct Xoshiro {
  uint64_t s[4];
  bool has_spare = false;
  double spare = 0;
  explicit Xoshiro(uint64_t seed) {
    SplitMix64 sm{seed};
    for (auto& v : s) v = sm.next();
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // [0, 1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  // integer in [0, m) by 128-bit multiply-shift (m < 2^32)
  uint64_t below(uint64_t m) { return (uint64_t)(((unsigned __int128)(next() >> 32) * m) >> 32); }
  // N(0,1) by the Box-Muller transform, both outputs used in turn
  double normal() {
    if (has_spare) {
      has_spare = false;
      return spare;
    }
    const double u1 = 1.0 - uniform();  // (0, 1]
    const double u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double a = 2.0 * M_PI * u2;
    spare = r * std::sin(a);
    has_spare = true;
    return r * std::cos(a);
  }
};

// Per-row stream: independent of thread count.
Xoshiro row_stream(uint64_t seed, int64_t row) {
  SplitMix64 sm{seed ^ ((uint64_t)row * 0xD1B54A32D192ED03ULL)};
  return Xoshiro(sm.next());
}

template <class F>
void parallel_rows(int64_t n, F f) {
  unsigned T = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
  if (n < 4096) T = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < T; ++t) {
    const int64_t lo = n * t / T, hi = n * (t + 1) / T;
    th.emplace_back([=]() {
      for (int64_t i = lo; i < hi; ++i) f(i);
    });
  }
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

int32_t ks_abi_version(void) { return KS_ABI_VERSION; }

// Reference defaults (solver_collection.cpp:20-49).
void ks_params_default(ks_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  std::strcpy(p->inner, "NO");
  p->outer_maxloop = 10000;
  p->outer_restart = 1000;
  p->outer_eps = 1e-8;
  p->outer_kskip = 2;
  p->outer_fix = 2;
  p->inner_maxloop = 50;
  p->inner_restart = 1000;
  p->inner_eps = 1e-1;
  p->inner_kskip = 2;
  p->inner_fix = 2;
  p->inner_precision = KS_F64;
  p->omp_threads = 8;
}

const char* ks_status_string(int32_t s) {
  switch (s) {
    case KS_OK: return "converge";
    case KS_NOT_CONVERGED: return "not converge";
    case KS_BREAKDOWN: return "breakdown";
    case KS_ERR_BAD_INPUT: return "bad input";
    case KS_ERR_HIP: return "HIP error";
    case KS_ERR_UNSUPPORTED: return "unsupported";
    case KS_ERR_NO_MEMORY: return "out of device memory";
    case KS_ERR_IO: return "I/O error";
    default: return "unknown status";
  }
}

ks_status ks_gen_vector(uint64_t seed, int32_t dist, double lo, double hi, int64_t n, double* out) {
  if (n < 0 || (n > 0 && !out)) return KS_ERR_BAD_INPUT;
  Xoshiro g(seed);
  if (dist == 0) {
    for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * g.uniform();
  } else if (dist == 1) {
    for (int64_t i = 0; i < n; ++i) out[i] = g.normal();
  } else {
    return KS_ERR_BAD_INPUT;
  }
  return KS_OK;
}

ks_status ks_gen_stencil(int32_t dim, int32_t nx, int32_t ny, int32_t nz, double pe, const double* vel,
                         int64_t* n_out, int64_t* nnz_out, int32_t* ptr, int32_t* col, double* val) {
  if (dim != 2 && dim != 3) return KS_ERR_BAD_INPUT;
  if (dim == 2) nz = 1;
  if (nx < 1 || ny < 1 || nz < 1) return KS_ERR_BAD_INPUT;
  const int64_t n = (int64_t)nx * ny * nz;
  int64_t nnz = n;
  nnz += 2 * ((int64_t)(nx - 1) * ny * nz + (int64_t)nx * (ny - 1) * nz);
  if (dim == 3) nnz += 2 * (int64_t)nx * ny * (nz - 1);
  if (n >= INT32_MAX || nnz >= INT32_MAX) return KS_ERR_BAD_INPUT;
  if (n_out) *n_out = n;
  if (nnz_out) *nnz_out = nnz;
  if (!ptr) return KS_OK;
  double c[3] = {0, 0, 0};
  if (pe != 0.0) {
    if (!vel) return KS_ERR_BAD_INPUT;
    double vmax = 0;
    for (int d = 0; d < dim; ++d) vmax = std::max(vmax, std::fabs(vel[d]));
    if (vmax == 0) return KS_ERR_BAD_INPUT;
    for (int d = 0; d < dim; ++d) c[d] = 2.0 * pe * vel[d] / vmax;
  }
  const double diag = 2.0 * dim + c[0] + c[1] + c[2];
  // row lengths, then fill in sorted column order: -z -y -x diag +x +y +z
  ptr[0] = 0;
  for (int64_t r = 0; r < n; ++r) {
    const int64_t i = r % nx, j = (r / nx) % ny, k = r / ((int64_t)nx * ny);
    int len = 1 + (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1);
    if (dim == 3) len += (k > 0) + (k < nz - 1);
    ptr[r + 1] = ptr[r] + len;
  }
  parallel_rows(n, [&](int64_t r) {
    const int64_t i = r % nx, j = (r / nx) % ny, k = r / ((int64_t)nx * ny);
    const int64_t sxy = (int64_t)nx * ny;
    int64_t q = ptr[r];
    if (dim == 3 && k > 0) { col[q] = (int32_t)(r - sxy); val[q] = -1.0 - c[2]; ++q; }
    if (j > 0) { col[q] = (int32_t)(r - nx); val[q] = -1.0 - c[1]; ++q; }
    if (i > 0) { col[q] = (int32_t)(r - 1); val[q] = -1.0 - c[0]; ++q; }
    col[q] = (int32_t)r; val[q] = diag; ++q;
    if (i < nx - 1) { col[q] = (int32_t)(r + 1); val[q] = -1.0; ++q; }
    if (j < ny - 1) { col[q] = (int32_t)(r + nx); val[q] = -1.0; ++q; }
    if (dim == 3 && k < nz - 1) { col[q] = (int32_t)(r + sxy); val[q] = -1.0; ++q; }
  });
  return KS_OK;
}

// Edge numbering: horizontal edge (i, j), i in [0, nx], joins (i-1, j) and
// (i, j) (i = 0 and i = nx are Dirichlet boundary edges): index j*(nx+1)+i.
// Vertical edge (i, j), j in [0, ny], joins (i, j-1) and (i, j): index
// (nx+1)*ny + j*nx + i.
ks_status ks_aniso2d_values(int32_t nx, int32_t ny, const double* kap, double* val) {
  if (nx < 1 || ny < 1 || !kap || !val) return KS_ERR_BAD_INPUT;
  const int64_t n = (int64_t)nx * ny;
  const int64_t nh = (int64_t)(nx + 1) * ny;
  auto kh = [&](int64_t i, int64_t j) { return kap[j * (nx + 1) + i]; };
  auto kv = [&](int64_t i, int64_t j) { return kap[nh + j * nx + i]; };
  // row lengths are the same as ks_gen_stencil(2, nx, ny)
  std::vector<int64_t> start(n + 1, 0);
  for (int64_t r = 0; r < n; ++r) {
    const int64_t i = r % nx, j = r / nx;
    start[r + 1] = start[r] + 1 + (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1);
  }
  parallel_rows(n, [&](int64_t r) {
    const int64_t i = r % nx, j = r / nx;
    int64_t q = start[r];
    if (j > 0) val[q++] = -kv(i, j);
    if (i > 0) val[q++] = -kh(i, j);
    val[q++] = kh(i, j) + kh(i + 1, j) + kv(i, j) + kv(i, j + 1);
    if (i < nx - 1) val[q++] = -kh(i + 1, j);
    if (j < ny - 1) val[q++] = -kv(i, j + 1);
  });
  return KS_OK;
}

ks_status ks_gen_aniso2d(int32_t nx, int32_t ny, uint64_t seed, int64_t* n_out, int64_t* nnz_out,
                         int32_t* ptr, int32_t* col, double* val, double* edge_kappa) {
  int64_t n, nnz;
  ks_status st = ks_gen_stencil(2, nx, ny, 1, 0.0, nullptr, &n, &nnz, nullptr, nullptr, nullptr);
  if (st != KS_OK) return st;
  if (n_out) *n_out = n;
  if (nnz_out) *nnz_out = nnz;
  if (!ptr) return KS_OK;
  st = ks_gen_stencil(2, nx, ny, 1, 0.0, nullptr, nullptr, nullptr, ptr, col, val);
  if (st != KS_OK) return st;
  const int64_t ne = (int64_t)(nx + 1) * ny + (int64_t)nx * (ny + 1);
  std::vector<double> kap(ne);
  Xoshiro g(seed);
  for (int64_t e = 0; e < ne; ++e) kap[e] = std::exp(g.normal());
  if (edge_kappa) std::memcpy(edge_kappa, kap.data(), sizeof(double) * ne);
  return ks_aniso2d_values(nx, ny, kap.data(), val);
}

ks_status ks_gen_zipf(int64_t n, double s, int32_t cap, uint64_t seed_len, uint64_t seed_col, uint64_t seed_val,
                      int64_t* nnz_out, int32_t* ptr, int32_t* col, double* val) {
  if (n < 2 || n >= INT32_MAX || cap < 1 || s <= 0) return KS_ERR_BAD_INPUT;
  if (cap > n) cap = (int32_t)n;
  // truncated Zipf on 1..cap: P(L = k) proportional to k^-s
  std::vector<double> cdf(cap);
  double acc = 0;
  for (int k = 1; k <= cap; ++k) {
    acc += std::pow((double)k, -s);
    cdf[k - 1] = acc;
  }
  for (auto& v : cdf) v /= acc;
  std::vector<int32_t> len(n);
  Xoshiro g(seed_len);
  int64_t nnz = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double u = g.uniform();
    const int k = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
    len[i] = std::min(k, cap);
    nnz += len[i];
  }
  if (nnz >= INT32_MAX) return KS_ERR_BAD_INPUT;
  if (nnz_out) *nnz_out = nnz;
  if (!ptr) return KS_OK;
  ptr[0] = 0;
  for (int64_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + len[i];
  parallel_rows(n, [&](int64_t i) {
    const int need = len[i] - 1;  // off-diagonal entries
    std::vector<int32_t> c;
    c.reserve(need + 8);
    Xoshiro gc = row_stream(seed_col, i);
    while ((int)c.size() < need) {
      while ((int)c.size() < need) {
        const int32_t v = (int32_t)gc.below((uint64_t)(n - 1));
        c.push_back(v >= i ? v + 1 : v);  // uniform over [0, n) \ {i}
      }
      std::sort(c.begin(), c.end());
      c.erase(std::unique(c.begin(), c.end()), c.end());
    }
    c.push_back((int32_t)i);
    std::sort(c.begin(), c.end());
    Xoshiro gv = row_stream(seed_val, i);
    const int64_t q0 = ptr[i];
    double asum = 0;
    int64_t dpos = -1;
    for (size_t t = 0; t < c.size(); ++t) {
      col[q0 + t] = c[t];
      if (c[t] == i) {
        dpos = q0 + (int64_t)t;
      } else {
        const double v = gv.normal();
        val[q0 + t] = v;
        asum += std::fabs(v);
      }
    }
    val[dpos] = need > 0 ? 1.1 * asum : 1.0;
  });
  return KS_OK;
}

}  // extern "C"

// ---------------------------------------------------------- CRS text ----
namespace {
bool read_all(const std::string& path, std::string& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(sz > 0 ? (size_t)sz : 0);
  size_t got = sz > 0 ? std::fread(&out[0], 1, (size_t)sz, f) : 0;
  std::fclose(f);
  return got == out.size();
}
bool header(const std::string& path, int64_t h[3]) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  long long a, b, c;
  int k = std::fscanf(f, "%lld %lld %lld", &a, &b, &c);
  std::fclose(f);
  if (k != 3) return false;
  h[0] = a; h[1] = b; h[2] = c;
  return true;
}

unsigned io_threads(size_t bytes) {
  if (bytes < (1u << 20)) return 1;
  return std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

inline bool is_ws(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Parses the first `ntok` whitespace-separated tokens of [b, e) with
// fn(token_index, p, &end) -> ok, in parallel: the text is cut at token
// starts, one pass counts each part's tokens, a second parses them at their
// global indices.  Any whitespace layout is accepted (as fscanf would); a
// token not followed by whitespace or the end of the text is an error.
template <class Fn>
bool parse_tokens(const char* b, const char* e, int64_t ntok, Fn fn) {
  const unsigned T = io_threads((size_t)(e - b));
  std::vector<const char*> cut(T + 1);
  cut[0] = b;
  cut[T] = e;
  for (unsigned t = 1; t < T; ++t) {
    const char* p = b + (e - b) * (int64_t)t / T;
    while (p < e && p > b && !is_ws(p[-1])) ++p;
    while (p < e && is_ws(*p)) ++p;
    cut[t] = std::max(p, cut[t - 1]);
  }
  std::vector<int64_t> cnt(T + 1, 0);
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t)
      th.emplace_back([&, t]() {
        int64_t c = 0;
        for (const char* q = cut[t]; q < cut[t + 1]; ++q)
          if (!is_ws(*q) && (q == cut[t] || is_ws(q[-1]))) ++c;
        cnt[t + 1] = c;
      });
    for (auto& x : th) x.join();
  }
  for (unsigned t = 0; t < T; ++t) cnt[t + 1] += cnt[t];
  if (cnt[T] < ntok) return false;
  std::vector<char> ok(T, 1);
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t)
      th.emplace_back([&, t]() {
        int64_t idx = cnt[t];
        const char* q = cut[t];
        const char* lim = cut[t + 1];
        while (idx < ntok) {
          while (q < lim && is_ws(*q)) ++q;
          if (q >= lim) break;
          char* end = nullptr;
          if (!fn(idx, q, &end) || end == q || (end < e && !is_ws(*end))) { ok[t] = 0; return; }
          q = end;
          ++idx;
        }
      });
    for (auto& x : th) x.join();
  }
  for (char v : ok)
    if (!v) return false;
  return true;
}

// One column index or row pointer: a decimal integer in [0, INT32_MAX].  A
// value outside that range is a malformed file, not a different matrix, so
// it is rejected rather than narrowed.
bool parse_index(const char* p, char** end, int32_t* out) {
  errno = 0;
  const long long v = std::strtoll(p, end, 10);
  if (errno == ERANGE || v < 0 || v > INT32_MAX) return false;
  *out = (int32_t)v;
  return true;
}

// the text after the three header integers
const char* after_header(const std::string& buf) {
  const char* p = buf.c_str();
  char* end = nullptr;
  for (int i = 0; i < 3; ++i) {
    std::strtoll(p, &end, 10);
    p = end;
  }
  return p;
}

// ------------------------------------------------ binary CSR sidecar ----
// Layout (little-endian): a 64-byte header, then ptr (int32, n+1), col
// (int32, nnz), val (f64, nnz), b (f64, n) and x (f64, n), each section
// starting on an 8-byte boundary.  The checksum covers the sections.
struct KsbHeader {
  char magic[8];
  uint32_t version;
  uint32_t flags;  // 1: b present, 2: x present
  int64_t n, nnz;
  uint64_t stamp;     // fingerprint of the CRS text files it was made from
  uint64_t checksum;  // of the payload sections
  uint64_t reserved[2];
};
static_assert(sizeof(KsbHeader) == 64, "header is 64 bytes");
constexpr char kKsbMagic[8] = {'K', 'S', 'C', 'S', 'R', 'B', 'I', 'N'};
constexpr uint32_t kKsbVersion = 1;

inline uint64_t mix64(uint64_t z) {  // SplitMix64 finaliser
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Hash of one section: 1 MiB blocks hashed word by word (in parallel), the
// block hashes folded in block order - the value does not depend on the
// thread count.
uint64_t section_hash(const void* data, size_t len) {
  constexpr size_t kBlk = 1 << 20;
  const size_t nb = (len + kBlk - 1) / kBlk;
  std::vector<uint64_t> hb(nb);
  const unsigned T = std::max<unsigned>(1, std::min<size_t>(io_threads(len), nb));
  auto work = [&](unsigned t) {
    for (size_t i = t; i < nb; i += T) {
      const unsigned char* p = (const unsigned char*)data + i * kBlk;
      const size_t m = std::min(kBlk, len - i * kBlk);
      uint64_t h = mix64(0x6b73637372ULL + i);
      size_t j = 0;
      for (; j + 8 <= m; j += 8) {
        uint64_t w;
        std::memcpy(&w, p + j, 8);
        h = mix64(h ^ w);
      }
      if (j < m) {
        uint64_t w = 0;
        std::memcpy(&w, p + j, m - j);
        h = mix64(h ^ w ^ ((uint64_t)(m - j) << 56));
      }
      hb[i] = h;
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 0; t < T; ++t) th.emplace_back(work, t);
  for (auto& x : th) x.join();
  uint64_t h = mix64(len);
  for (uint64_t v : hb) h = mix64(h ^ v);
  return h;
}

struct Section {
  const void* p;
  size_t len;
};
uint64_t payload_hash(const Section* s, int k) {
  uint64_t h = 0x5eedULL;
  for (int i = 0; i < k; ++i) h = mix64(h ^ section_hash(s[i].p, s[i].len) ^ (uint64_t)i);
  return h;
}
inline size_t pad8(size_t b) { return (8 - b % 8) % 8; }
}  // namespace

extern "C" {

// collection::checkCRSMatrix (solver_collection.cpp:655-724): the three
// headers must agree on N, M, NNZ and N must equal M.
ks_status ks_crs_read_header(const char* dir, int64_t* n, int64_t* nnz) {
  if (!dir) return KS_ERR_BAD_INPUT;
  const std::string d(dir);
  int64_t h[3][3];
  const char* names[3] = {"/ColVal.txt", "/Ptr.txt", "/bx.txt"};
  for (int i = 0; i < 3; ++i)
    if (!header(d + names[i], h[i])) return KS_ERR_IO;
  for (int f = 0; f < 3; ++f)
    if (h[0][f] != h[1][f] || h[1][f] != h[2][f]) return KS_ERR_IO;
  if (h[0][0] != h[0][1]) return KS_ERR_IO;
  if (n) *n = h[0][0];
  if (nnz) *nnz = h[0][2];
  return KS_OK;
}

// collection::readMatrix (solver_collection.cpp:726-828): ColVal.txt holds
// nnz "col val" records, Ptr.txt n+1 row pointers, bx.txt n "b x" records,
// each after the "N M NNZ" header.  Parsed in parallel (parse_tokens).
ks_status ks_crs_read(const char* dir, int32_t* ptr, int32_t* col, double* val, double* b, double* x) {
  int64_t n, nnz;
  ks_status st = ks_crs_read_header(dir, &n, &nnz);
  if (st != KS_OK) return st;
  const std::string d(dir);
  std::string buf;
  if (!read_all(d + "/ColVal.txt", buf)) return KS_ERR_IO;
  const char* e = buf.c_str() + buf.size();
  bool ok = parse_tokens(after_header(buf), e, 2 * nnz, [&](int64_t i, const char* p, char** end) {
    if (i % 2 == 0) {
      int32_t c;
      if (!parse_index(p, end, &c)) return false;
      if (col) col[i / 2] = c;
    } else {
      const double v = std::strtod(p, end);
      if (val) val[i / 2] = v;
    }
    return true;
  });
  if (!ok) return KS_ERR_IO;
  if (!read_all(d + "/Ptr.txt", buf)) return KS_ERR_IO;
  e = buf.c_str() + buf.size();
  ok = parse_tokens(after_header(buf), e, n + 1, [&](int64_t i, const char* p, char** end) {
    int32_t c;
    if (!parse_index(p, end, &c)) return false;
    if (ptr) ptr[i] = c;
    return true;
  });
  if (!ok) return KS_ERR_IO;
  if (!read_all(d + "/bx.txt", buf)) return KS_ERR_IO;
  e = buf.c_str() + buf.size();
  ok = parse_tokens(after_header(buf), e, 2 * n, [&](int64_t i, const char* p, char** end) {
    const double v = std::strtod(p, end);
    if (i % 2 == 0) {
      if (b) b[i / 2] = v;
    } else if (x) {
      x[i / 2] = v;
    }
    return true;
  });
  return ok ? KS_OK : KS_ERR_IO;
}

ks_status ks_crs_write(const char* dir, int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col,
                       const double* val, const double* b, const double* x) {
  if (!dir || !ptr || (nnz > 0 && (!col || !val))) return KS_ERR_BAD_INPUT;
  const std::string d(dir);
  FILE* f = std::fopen((d + "/ColVal.txt").c_str(), "w");
  if (!f) return KS_ERR_IO;
  std::fprintf(f, "%lld %lld %lld\n", (long long)n, (long long)n, (long long)nnz);
  for (int64_t z = 0; z < nnz; ++z) std::fprintf(f, "%d %.17e\n", col[z], val[z]);
  std::fclose(f);
  f = std::fopen((d + "/Ptr.txt").c_str(), "w");
  if (!f) return KS_ERR_IO;
  std::fprintf(f, "%lld %lld %lld\n", (long long)n, (long long)n, (long long)nnz);
  for (int64_t z = 0; z < n + 1; ++z) std::fprintf(f, "%d\n", ptr[z]);
  std::fclose(f);
  f = std::fopen((d + "/bx.txt").c_str(), "w");
  if (!f) return KS_ERR_IO;
  std::fprintf(f, "%lld %lld %lld\n", (long long)n, (long long)n, (long long)nnz);
  for (int64_t z = 0; z < n; ++z) std::fprintf(f, "%.17e %.17e\n", b ? b[z] : 0.0, x ? x[z] : 0.0);
  std::fclose(f);
  return KS_OK;
}

// Fingerprint (size and modification time) of the three CRS text files.
ks_status ks_crs_stamp(const char* dir, uint64_t* stamp) {
  if (!dir || !stamp) return KS_ERR_BAD_INPUT;
  const std::string d(dir);
  uint64_t h = 0x63727374616d70ULL;
  for (const char* nm : {"/ColVal.txt", "/Ptr.txt", "/bx.txt"}) {
    struct stat sb;
    if (::stat((d + nm).c_str(), &sb) != 0) return KS_ERR_IO;
    h = mix64(h ^ (uint64_t)sb.st_size);
    h = mix64(h ^ (uint64_t)sb.st_mtim.tv_sec);
    h = mix64(h ^ (uint64_t)sb.st_mtim.tv_nsec);
  }
  *stamp = h ? h : 1;
  return KS_OK;
}

ks_status ks_crsbin_write(const char* path, int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col,
                          const double* val, const double* b, const double* x, uint64_t stamp) {
  if (!path || n < 1 || nnz < 0 || !ptr || (nnz > 0 && (!col || !val))) return KS_ERR_BAD_INPUT;
  KsbHeader h;
  std::memset(&h, 0, sizeof h);
  std::memcpy(h.magic, kKsbMagic, 8);
  h.version = kKsbVersion;
  h.flags = (b ? 1u : 0u) | (x ? 2u : 0u);
  h.n = n;
  h.nnz = nnz;
  h.stamp = stamp;
  Section sec[5] = {{ptr, sizeof(int32_t) * (size_t)(n + 1)}, {col, sizeof(int32_t) * (size_t)nnz},
                    {val, sizeof(double) * (size_t)nnz}, {b, b ? sizeof(double) * (size_t)n : 0},
                    {x, x ? sizeof(double) * (size_t)n : 0}};
  h.checksum = payload_hash(sec, 5);
  // written to a temporary name, then renamed: a reader never sees a partial file
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return KS_ERR_IO;
  static const char zeros[8] = {0};
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  for (int i = 0; i < 5 && ok; ++i) {
    if (sec[i].len) ok = std::fwrite(sec[i].p, 1, sec[i].len, f) == sec[i].len;
    const size_t pd = pad8(sec[i].len);
    if (ok && pd) ok = std::fwrite(zeros, 1, pd, f) == pd;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return KS_ERR_IO;
  }
  return KS_OK;
}

}  // extern "C"

namespace {
ks_status ksb_header(FILE* f, KsbHeader& h) {
  if (std::fread(&h, sizeof h, 1, f) != 1) return KS_ERR_IO;
  if (std::memcmp(h.magic, kKsbMagic, 8) != 0 || h.version != kKsbVersion) return KS_ERR_IO;
  if (h.n < 1 || h.n >= (int64_t)INT32_MAX || h.nnz < 0 || h.nnz >= (int64_t)INT32_MAX) return KS_ERR_IO;
  return KS_OK;
}
}  // namespace

extern "C" {

ks_status ks_crsbin_read_header(const char* path, int64_t* n, int64_t* nnz, uint64_t* stamp) {
  if (!path) return KS_ERR_BAD_INPUT;
  FILE* f = std::fopen(path, "rb");
  if (!f) return KS_ERR_IO;
  KsbHeader h;
  const ks_status st = ksb_header(f, h);
  std::fclose(f);
  if (st != KS_OK) return st;
  if (n) *n = h.n;
  if (nnz) *nnz = h.nnz;
  if (stamp) *stamp = h.stamp;
  return KS_OK;
}

// Reads the sections into the given arrays (NULL: read and checked, not
// returned); b or x absent from the file read as zeros.  The file size and
// the checksum must match the header.
ks_status ks_crsbin_read(const char* path, int32_t* ptr, int32_t* col, double* val, double* b, double* x) {
  if (!path) return KS_ERR_BAD_INPUT;
  FILE* f = std::fopen(path, "rb");
  if (!f) return KS_ERR_IO;
  KsbHeader h;
  ks_status st = ksb_header(f, h);
  if (st != KS_OK) {
    std::fclose(f);
    return st;
  }
  const size_t n = (size_t)h.n, nnz = (size_t)h.nnz;
  const bool hb = h.flags & 1u, hx = h.flags & 2u;
  size_t len[5] = {4 * (n + 1), 4 * nnz, 8 * nnz, hb ? 8 * n : 0, hx ? 8 * n : 0};
  size_t expect = sizeof h;
  for (size_t l : len) expect += l + pad8(l);
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  if (sz < 0 || (size_t)sz != expect) {
    std::fclose(f);
    return KS_ERR_IO;
  }
  std::fseek(f, sizeof h, SEEK_SET);
  void* dst[5] = {ptr, col, val, b, x};
  std::vector<std::vector<char>> tmp(5);
  Section sec[5];
  bool ok = true;
  for (int i = 0; i < 5 && ok; ++i) {
    void* p = dst[i];
    if (!p || !len[i]) {
      tmp[i].resize(len[i]);
      p = tmp[i].data();
    }
    if (len[i]) ok = std::fread(p, 1, len[i], f) == len[i];
    char padb[8];
    const size_t pd = pad8(len[i]);
    if (ok && pd) ok = std::fread(padb, 1, pd, f) == pd;
    sec[i] = Section{p, len[i]};
  }
  std::fclose(f);
  if (!ok || payload_hash(sec, 5) != h.checksum) return KS_ERR_IO;
  if (b && !hb) std::fill(b, b + n, 0.0);
  if (x && !hx) std::fill(x, x + n, 0.0);
  return KS_OK;
}

}  // extern "C"
