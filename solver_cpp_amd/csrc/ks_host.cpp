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
}  // namespace

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

// collection::readMatrix (solver_collection.cpp:726-828)
ks_status ks_crs_read(const char* dir, int32_t* ptr, int32_t* col, double* val, double* b, double* x) {
  int64_t n, nnz;
  ks_status st = ks_crs_read_header(dir, &n, &nnz);
  if (st != KS_OK) return st;
  const std::string d(dir);
  std::string buf;
  auto skip_header = [](const char*& p, char** end) {
    for (int i = 0; i < 3; ++i) { std::strtoll(p, end, 10); p = *end; }
  };
  char* end = nullptr;
  if (!read_all(d + "/ColVal.txt", buf)) return KS_ERR_IO;
  const char* p = buf.c_str();
  skip_header(p, &end);
  for (int64_t z = 0; z < nnz; ++z) {
    long c = std::strtol(p, &end, 10);
    if (end == p) return KS_ERR_IO;
    p = end;
    double v = std::strtod(p, &end);
    if (end == p) return KS_ERR_IO;
    p = end;
    if (col) col[z] = (int32_t)c;
    if (val) val[z] = v;
  }
  if (!read_all(d + "/Ptr.txt", buf)) return KS_ERR_IO;
  p = buf.c_str();
  skip_header(p, &end);
  for (int64_t z = 0; z < n + 1; ++z) {
    long c = std::strtol(p, &end, 10);
    if (end == p) return KS_ERR_IO;
    p = end;
    if (ptr) ptr[z] = (int32_t)c;
  }
  if (!read_all(d + "/bx.txt", buf)) return KS_ERR_IO;
  p = buf.c_str();
  skip_header(p, &end);
  for (int64_t z = 0; z < n; ++z) {
    double v1 = std::strtod(p, &end);
    if (end == p) return KS_ERR_IO;
    p = end;
    double v2 = std::strtod(p, &end);
    if (end == p) return KS_ERR_IO;
    p = end;
    if (b) b[z] = v1;
    if (x) x[z] = v2;
  }
  return KS_OK;
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

}  // extern "C"
