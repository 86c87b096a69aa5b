// ks_device.hip - libksolve entry points that touch the GPU: context, matrix
// (upload, FP32 copy, device transpose), BLAS-layer operators, ks_solve and
// ks_adjoint.  All work is ordered on the context's HIP stream.
#include <atomic>
#include <chrono>
#include <memory>
#include <mutex>
#include <thread>

#include "ks_levels.h"

using namespace ks;

namespace {

enum Name { NONE, CG, BICG, CR, GCR, GMRES, KSKIPCG, KSKIPCR, KSKIPBICG, VPCG, VPBICG, VPCR, VPGCR, VPGMRES };

// collection::string2enum (solver_collection.cpp:238-267)
Name name_of(const char* s16) {
  char buf[17];
  std::memcpy(buf, s16, 16);
  buf[16] = 0;
  const std::string t(buf);
  static const char* lo[] = {"", "cg", "bicg", "cr", "gcr", "gmres", "kskipcg", "kskipcr", "kskipbicg",
                             "vpcg", "vpbicg", "vpcr", "vpgcr", "vpgmres"};
  static const char* up[] = {"", "CG", "BICG", "CR", "GCR", "GMRES", "KSKIPCG", "KSKIPCR", "KSKIPBICG",
                             "VPCG", "VPBICG", "VPCR", "VPGCR", "VPGMRES"};
  for (int i = 1; i <= 13; ++i)
    if (t == lo[i] || t == up[i]) return (Name)i;
  return NONE;
}

// Host or device input, uploaded when host.
template <class T>
struct In {
  const T* p = nullptr;
  DBuf<T> own;
  In(ks_ctx* c, const T* src, int64_t n, uint32_t flags) {
    if ((flags & KS_PTR_DEVICE) || src == nullptr || n == 0) {
      p = src;
    } else {
      own.alloc(n);
      KS_HIPCK(hipMemcpyAsync(own.p, src, sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
      p = own.p;
    }
  }
};
// Output (device or host); copy_back() after the work.
template <class T>
struct Out {
  T* p = nullptr;
  T* host = nullptr;
  int64_t n = 0;
  DBuf<T> own;
  Out(ks_ctx* c, T* dst, int64_t n_, uint32_t flags, bool copy_in = false) : n(n_) {
    if ((flags & KS_PTR_DEVICE) || dst == nullptr || n == 0) {
      p = dst;
    } else {
      own.alloc(n);
      host = dst;
      p = own.p;
      if (copy_in) KS_HIPCK(hipMemcpyAsync(own.p, dst, sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
    }
  }
  void copy_back(ks_ctx* c) {
    if (host) KS_HIPCK(hipMemcpyAsync(host, own.p, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
  }
};

void set_device(ks_ctx* c) { KS_HIPCK(hipSetDevice(c->device)); }

hipEvent_t ev_get(ks_ctx* c) {
  if (c->evpool.empty()) {
    hipEvent_t e;
    KS_HIPCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
  }
  hipEvent_t e = c->evpool.back();
  c->evpool.pop_back();
  return e;
}

// Per-element dot / norm / axpy ops for the BLAS entry points.
template <class T>
struct OpDot {
  const T *x, *y;
  double* out;
  int sq;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double* acc) { acc[0] += (double)x[i] * (double)y[i]; }
  __device__ void fin(const double* s) { *out = sq ? sqrt(s[0]) : s[0]; }
};
struct OpAxpyApi {
  double a;
  const double *x, *y;
  double* out;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double*) { out[i] = a * x[i] + y[i]; }
  __device__ void fin(const double*) {}
};
struct OpResid2 {  // (r,r) of an existing residual
  const double* r;
  double* out;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double* acc) { acc[0] += r[i] * r[i]; }
  __device__ void fin(const double* s) { *out = s[0]; }
};

// ---------------------------------------------------------- transpose ---
// CSR-stream tile sets requested by the matrix flags (make_plan's `tiles`).
int tile_choice(uint32_t flags) {
  if (flags & KS_MAT_WG_TILES) return 1;
  if (flags & KS_MAT_WAVE_TILES) return 2;
  return 0;
}

// Merge path for KS_MAT_MERGE, and by default (no layout flag) for
// heavy-tailed row lengths: a longest row above 256 entries and 16 times the
// mean (the Zipf-like case, where it beats CSR-stream tiles).
bool merge_choice(uint32_t flags, int64_t n, int64_t nnz, int64_t max_row) {
  if (flags & KS_MAT_MERGE) return true;
  if (flags & (KS_MAT_VECTOR | KS_MAT_WG_TILES | KS_MAT_WAVE_TILES)) return false;
  const double mean = n > 0 ? (double)nnz / (double)n : 0.0;
  return max_row > 256 && (double)max_row > 16.0 * mean;
}

// Column panels (PanelSet): with KS_MAT_PANELS, and by default (no layout
// flag) for a matrix with at least 2^21 rows whose gathers are mostly far
// from the diagonal - over half of the entries of sampled rows more than n/8
// columns from their row, as with random columns or a scrambled ordering (a
// banded matrix, or one whose rows add a few far columns to a local band,
// gathers mostly near its rows and keeps one pass) - the SpMV runs as kPanels
// passes over column slices of x small enough for an XCD's 4 MB L2 (n / 10
// entries: 3.2 MB at n = 4M in FP64).  Ten is the fastest count measured on
// the 4M-row Zipf matrix (FP64 1.89 ms against 1.96 / 1.93 / 1.91 / 2.07 ms
// with 6 / 8 / 12 / 16 panels): fewer passes miss L2 more on the gathers,
// more passes re-read every row's pointer and partial sum more often.
constexpr int kPanels = 10;
bool panel_choice(uint32_t flags, int64_t n, const int32_t* hp, const int32_t* hc) {
  if (flags & KS_MAT_PANELS) return true;
  if (flags & (KS_MAT_VECTOR | KS_MAT_WG_TILES | KS_MAT_WAVE_TILES | KS_MAT_MERGE)) return false;
  if (n < (int64_t(1) << 21)) return false;
  int64_t far = 0, all = 0;
  for (int64_t r = 0; r < n; r += 61)
    for (int32_t j = hp[r]; j < hp[r + 1]; ++j) {
      far += std::llabs((int64_t)hc[j] - r) > n / 8;
      ++all;
    }
  return all > 0 && 2 * far > all;
}

// Panel values from A's values (and their FP32 copy when there is one).
void refresh_panels(ks_mat* m, PanelSet& ps) {
  if (!ps.np) return;
  ks_ctx* c = m->ctx;
  const int64_t tot = ps.off[ps.np];
  k_gather<double, double><<<grid_map(tot), kBlock, 0, c->stream>>>(tot, ps.perm.p, m->val.p, ps.val.p);
  KS_HIPCK(hipGetLastError());
  if (ps.val32.p) {
    k_convert<double, float><<<grid_map(tot), kBlock, 0, c->stream>>>(tot, ps.val.p, ps.val32.p, Gate{});
    KS_HIPCK(hipGetLastError());
  }
}

// FP32 panel values and views (after the matrix has its FP32 copy).
void panels_fp32(ks_mat* m, PanelSet& ps) {
  if (!ps.np || ps.val32.p) return;
  ks_ctx* c = m->ctx;
  const int64_t tot = ps.off[ps.np];
  ps.val32.alloc(tot);
  ps.acc32.alloc(m->n);
  k_convert<double, float><<<grid_map(tot), kBlock, 0, c->stream>>>(tot, ps.val.p, ps.val32.p, Gate{});
  KS_HIPCK(hipGetLastError());
  ps.v32.assign(ps.np, MatView<float>{});
  for (int p = 0; p < ps.np; ++p) {
    MatView<float>& v = ps.v32[p];
    v.n = m->n; v.nnz = ps.v64[p].nnz;
    v.ptr = ps.v64[p].ptr; v.col = ps.v64[p].col; v.val = ps.val32.p + ps.off[p];
    v.set_plan(ps.plans[p]);
  }
}

// fn(lo, hi) over [0, n) in contiguous row blocks, on up to 16 host threads.
template <class F>
void host_rows(int64_t n, F fn) {
  const int64_t T = std::min<int64_t>(16, std::max(1u, std::thread::hardware_concurrency()));
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(T, n / 65536));
  if (nt == 1) {
    fn(int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  for (int64_t t = 0; t < nt; ++t) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
  for (auto& x : th) x.join();
}

// 16-bit column offsets: the workgroup CSR-stream SpMV reads col - row as an
// int16 instead of the int32 column (2 bytes per entry less; stencil
// matrices), unless KS_MAT_COL32.  Every offset must lie in [-32767, 32767]:
// A^T's offsets are A's negated, so both fit.
bool offsets_fit16(int64_t n, const int32_t* hp, const int32_t* hc) {
  std::atomic<bool> ok{true};
  host_rows(n, [&](int64_t lo, int64_t hi) {
    for (int64_t r = lo; r < hi && ok.load(std::memory_order_relaxed); ++r)
      for (int32_t j = hp[r]; j < hp[r + 1]; ++j)
        if (std::llabs((int64_t)hc[j] - r) > 32767) {
          ok = false;
          break;
        }
  });
  return ok;
}

void build_col16(ks_mat* m, DBuf<int16_t>& out, const int32_t* ptr, const int32_t* col) {
  ks_ctx* c = m->ctx;
  out.alloc(m->nnz);
  k_col_offsets<<<grid_map(m->n), kBlock, 0, c->stream>>>(m->n, ptr, col, out.p);
  KS_HIPCK(hipGetLastError());
}

// Delta-coded 16-bit columns (k_spmv_stream's uint16_t format) of the CSR
// (ptr, col) with plan pl, when its workgroup tiles sum one row per lane and
// every row encodes (sorted, first column and gaps in range: checked on the
// device, as A^T's columns exist only there).  Returns false, with `out`
// empty, otherwise.
bool build_col_delta(ks_mat* m, const Plan& pl, DBuf<uint16_t>& out, const int32_t* ptr, const int32_t* col) {
  if (pl.kind != 0 || !(pl.wg64 || pl.wg32) || !pl.wg.lane1) return false;
  ks_ctx* c = m->ctx;
  out.alloc(m->nnz);
  DBuf<int> bad(1);
  KS_HIPCK(hipMemsetAsync(bad.p, 0, sizeof(int), c->stream));
  k_col_delta<<<grid_map(m->n), kBlock, 0, c->stream>>>(m->n, ptr, col, out.p, bad.p);
  KS_HIPCK(hipGetLastError());
  int hb = 0;
  KS_HIPCK(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  KS_HIPCK(hipStreamSynchronize(c->stream));
  if (hb) out.release();
  return !hb;
}

// The SpMV's 16-bit column indices of A and A^T: int16 offsets when every
// offset fits (off16), delta-coded columns where they apply, in the order
// KS_MAT_COL_DELTA asks for; int32 columns otherwise, and for a side whose
// SpMV runs on column panels (which read int32 columns).
void build_index16(ks_mat* m, bool off16) {
  const bool delta_first = (m->flags & KS_MAT_COL_DELTA) != 0;
  auto side = [&](DBuf<int16_t>& o16, DBuf<uint16_t>& od, const Plan& pl, const PanelSet& ps, const int32_t* ptr,
                  const int32_t* col) {
    if (ps.np) return;
    if (off16 && !delta_first) build_col16(m, o16, ptr, col);
    else if (!build_col_delta(m, pl, od, ptr, col) && off16) build_col16(m, o16, ptr, col);
  };
  side(m->col16, m->cold, m->plan, m->panels, m->ptr.p, m->col.p);
  if (m->has_t) side(m->tcol16, m->tcold, m->tplan, m->tpanels, m->tptr.p, m->tcol.p);
}

// Split the CSR (hp, hc: host row pointers and columns of A or A^T) into
// column panels.  src_perm (device) maps the CSR's entries to A's entries
// (A^T: tperm), or is null for A itself.
void build_panels(ks_mat* m, PanelSet& ps, const int32_t* hp, const int32_t* hc, const int32_t* src_perm,
                  uint32_t flags) {
  ks_ctx* c = m->ctx;
  const int64_t n = m->n;
  const int P = kPanels;
  auto panel_of = [&](int32_t col) { return (int)((int64_t)col * P / n); };
  // per-row counts (each row written by one thread), then per-panel scans
  std::vector<int32_t> hpp((size_t)P * (n + 1), 0);
  host_rows(n, [&](int64_t lo, int64_t hi) {
    for (int64_t r = lo; r < hi; ++r)
      for (int32_t j = hp[r]; j < hp[r + 1]; ++j) hpp[(size_t)panel_of(hc[j]) * (n + 1) + r + 1]++;
  });
  std::vector<int64_t> maxrow(P, 0);
  {
    std::vector<std::thread> th;
    for (int p = 0; p < P; ++p)
      th.emplace_back([&, p] {
        int32_t* q = hpp.data() + (size_t)p * (n + 1);
        for (int64_t r = 0; r < n; ++r) {
          maxrow[p] = std::max<int64_t>(maxrow[p], q[r + 1]);
          q[r + 1] += q[r];
        }
      });
    for (auto& x : th) x.join();
  }
  std::vector<int64_t> off(P + 1, 0);
  for (int p = 0; p < P; ++p)
    off[p + 1] = (off[p] + hpp[(size_t)p * (n + 1) + n] + 3) / 4 * 4;  // 16-byte aligned panel starts
  const int64_t tot = std::max<int64_t>(off[P], 1);
  std::vector<int32_t> perm(tot, 0), pcol(tot, 0);
  // each row's entries go to their panels in CSR order (rows are disjoint)
  host_rows(n, [&](int64_t lo, int64_t hi) {
    std::vector<int64_t> pos(P);
    for (int64_t r = lo; r < hi; ++r) {
      for (int p = 0; p < P; ++p) pos[p] = off[p] + hpp[(size_t)p * (n + 1) + r];
      for (int32_t j = hp[r]; j < hp[r + 1]; ++j) {
        const int p = panel_of(hc[j]);
        perm[pos[p]] = j;
        pcol[pos[p]] = hc[j];
        ++pos[p];
      }
    }
  });
  hipStream_t s = c->stream;
  ps.ptr.alloc(hpp.size());
  ps.col.alloc(tot);
  ps.perm.alloc(tot);
  ps.val.alloc(tot);
  ps.acc64.alloc(n);
  KS_HIPCK(hipMemcpyAsync(ps.ptr.p, hpp.data(), sizeof(int32_t) * hpp.size(), hipMemcpyHostToDevice, s));
  KS_HIPCK(hipMemcpyAsync(ps.col.p, pcol.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice, s));
  if (src_perm) {
    DBuf<int32_t> tmp(tot);
    KS_HIPCK(hipMemcpyAsync(tmp.p, perm.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice, s));
    k_gather<int32_t, int32_t><<<grid_map(tot), kBlock, 0, s>>>(tot, tmp.p, src_perm, ps.perm.p);
    KS_HIPCK(hipGetLastError());
    KS_HIPCK(hipStreamSynchronize(s));
  } else {
    KS_HIPCK(hipMemcpyAsync(ps.perm.p, perm.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice, s));
  }
  ps.off = off;
  ps.plans.clear();
  ps.plans.resize(P);
  ps.v64.assign(P, MatView<double>{});
  const uint32_t lay = flags & ~(uint32_t)KS_MAT_PANELS;
  for (int p = 0; p < P; ++p) {
    const int32_t* q = hpp.data() + (size_t)p * (n + 1);
    const int64_t np_nnz = q[n];
    make_plan(c, ps.plans[p], q, n, np_nnz, (lay & KS_MAT_VECTOR) != 0, tile_choice(lay),
              merge_choice(lay, n, np_nnz, maxrow[p]));
    MatView<double>& v = ps.v64[p];
    v.n = n; v.nnz = np_nnz;
    v.ptr = ps.ptr.p + (size_t)p * (n + 1); v.col = ps.col.p + off[p]; v.val = ps.val.p + off[p];
    v.set_plan(ps.plans[p]);
  }
  ps.np = P;
  refresh_panels(m, ps);
  KS_HIPCK(hipStreamSynchronize(s));  // the host arrays above are freed on return
}

// A^T row segments longer than kSortMax: bitonic sorts of kSortMax-entry
// chunks, then pairwise merges of sorted runs (width kSortMax, 2 kSortMax,
// ...) that alternate between A^T's arrays and a flat staging buffer.
void sort_huge_segments(ks_mat* m, const std::vector<int32_t>& htp, hipStream_t s) {
  const int64_t n = m->n;
  std::vector<int32_t> cs, cl;
  std::vector<int64_t> off{0}, sb, fb;
  int64_t longest = 0;
  for (int64_t r = 0; r < n; ++r) {
    const int64_t a = htp[r], len = htp[r + 1] - a;
    if (len <= kSortMax) continue;
    for (int64_t q = 0; q < len; q += kSortMax) {
      cs.push_back((int32_t)(a + q));
      cl.push_back((int32_t)std::min<int64_t>(kSortMax, len - q));
    }
    sb.push_back(a);
    fb.push_back(off.back());
    off.push_back(off.back() + len);
    longest = std::max(longest, len);
  }
  const int32_t nh = (int32_t)sb.size();
  const int64_t total = off.back();
  DBuf<int32_t> dcs(cs.size()), dcl(cl.size()), fc(total), fp(total);
  DBuf<int64_t> doff(off.size()), dsb(nh), dfb(nh);
  KS_HIPCK(hipMemcpyAsync(dcs.p, cs.data(), sizeof(int32_t) * cs.size(), hipMemcpyHostToDevice, s));
  KS_HIPCK(hipMemcpyAsync(dcl.p, cl.data(), sizeof(int32_t) * cl.size(), hipMemcpyHostToDevice, s));
  KS_HIPCK(hipMemcpyAsync(doff.p, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice, s));
  KS_HIPCK(hipMemcpyAsync(dsb.p, sb.data(), sizeof(int64_t) * nh, hipMemcpyHostToDevice, s));
  KS_HIPCK(hipMemcpyAsync(dfb.p, fb.data(), sizeof(int64_t) * nh, hipMemcpyHostToDevice, s));
  k_sort_chunks<<<(unsigned)cs.size(), 1024, 0, s>>>(dcs.p, dcl.p, m->tcol.p, m->tperm.p);
  KS_HIPCK(hipGetLastError());
  bool in_flat = false;
  for (int64_t w = kSortMax; w < longest; w *= 2) {
    if (in_flat)
      k_merge_runs<<<grid_map(total), kBlock, 0, s>>>(total, nh, doff.p, dfb.p, dsb.p, w, fc.p, fp.p, m->tcol.p,
                                                      m->tperm.p);
    else
      k_merge_runs<<<grid_map(total), kBlock, 0, s>>>(total, nh, doff.p, dsb.p, dfb.p, w, m->tcol.p, m->tperm.p,
                                                      fc.p, fp.p);
    KS_HIPCK(hipGetLastError());
    in_flat = !in_flat;
  }
  if (in_flat) {  // a run as wide as every segment has no partner: a plain copy back
    k_merge_runs<<<grid_map(total), kBlock, 0, s>>>(total, nh, doff.p, dfb.p, dsb.p, int64_t(1) << 40, fc.p, fp.p,
                                                    m->tcol.p, m->tperm.p);
    KS_HIPCK(hipGetLastError());
  }
  KS_HIPCK(hipStreamSynchronize(s));
}

void build_transpose(ks_mat* m) {
  ks_ctx* c = m->ctx;
  const int64_t n = m->n, nnz = m->nnz;
  hipStream_t s = c->stream;
  DBuf<int32_t> cnt(n + 1), fill(n + 1);
  m->tptr.alloc(n + 1);
  m->tcol.alloc(std::max<int64_t>(nnz, 1));
  m->tperm.alloc(std::max<int64_t>(nnz, 1));
  m->tval.alloc(std::max<int64_t>(nnz, 1));
  KS_HIPCK(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (n + 1), s));
  KS_HIPCK(hipMemsetAsync(fill.p, 0, sizeof(int32_t) * (n + 1), s));
  if (nnz > 0) {
    k_col_count<<<grid_map(nnz), kBlock, 0, s>>>(nnz, m->col.p, cnt.p);
    KS_HIPCK(hipGetLastError());
  }
  const int64_t nb = (n + 1023) / 1024;
  DBuf<int64_t> bsum(nb + 1);
  k_scan_blocks<<<nb, kBlock, 0, s>>>(n, cnt.p, bsum.p);
  k_scan_serial<<<1, 1, 0, s>>>(nb, bsum.p);
  k_scan_apply<<<nb, kBlock, 0, s>>>(n, cnt.p, bsum.p, m->tptr.p);
  KS_HIPCK(hipGetLastError());
  if (nnz > 0) {
    k_tscatter<8><<<grid_map(n), kBlock, 0, s>>>(n, m->ptr.p, m->col.p, m->tptr.p, fill.p, m->tcol.p, m->tperm.p);
    KS_HIPCK(hipGetLastError());
  }
  // per-segment sort makes the result independent of the scatter order
  DBuf<int32_t> longrows(n + 1), midrows(n + 1);
  KS_HIPCK(hipMemsetAsync(longrows.p + n, 0, sizeof(int32_t), s));
  KS_HIPCK(hipMemsetAsync(midrows.p + n, 0, sizeof(int32_t), s));
  k_sort_small<<<grid_map(n), kBlock, 0, s>>>(n, m->tptr.p, m->tcol.p, m->tperm.p, longrows.p, longrows.p + n,
                                              midrows.p, midrows.p + n);
  KS_HIPCK(hipGetLastError());
  std::vector<int32_t> htp(n + 1);
  int32_t nlong = 0, nmid = 0;
  KS_HIPCK(hipMemcpyAsync(htp.data(), m->tptr.p, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, s));
  KS_HIPCK(hipMemcpyAsync(&nlong, longrows.p + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  KS_HIPCK(hipMemcpyAsync(&nmid, midrows.p + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  KS_HIPCK(hipStreamSynchronize(s));
  KS_CHECK(htp[n] == nnz, KS_ERR_BAD_INPUT, "transpose: column counts do not add up to nnz");
  int64_t mx = 0;
  for (int64_t i = 0; i < n; ++i) mx = std::max<int64_t>(mx, htp[i + 1] - htp[i]);
  m->max_trow = mx;
  if (nmid > 0) {
    k_sort_mid<<<(nmid + kBlock / 64 - 1) / (kBlock / 64), kBlock, 0, s>>>(nmid, m->tptr.p, midrows.p, m->tcol.p,
                                                                         m->tperm.p);
    KS_HIPCK(hipGetLastError());
  }
  if (nlong > 0) {
    k_sort_long<<<nlong, 1024, 0, s>>>(m->tptr.p, longrows.p, m->tcol.p, m->tperm.p);
    KS_HIPCK(hipGetLastError());
  }
  if (mx > kSortMax) sort_huge_segments(m, htp, s);
  if (nnz > 0) {
    k_gather<double, double><<<grid_map(nnz), kBlock, 0, s>>>(nnz, m->tperm.p, m->val.p, m->tval.p);
    KS_HIPCK(hipGetLastError());
  }
  make_plan(c, m->tplan, htp.data(), n, nnz, (m->flags & KS_MAT_VECTOR) != 0, tile_choice(m->flags),
            merge_choice(m->flags, n, nnz, mx));
  m->has_t = true;
  if (m->panels.np && nnz > 0) {  // A^T gathers as randomly as A
    std::vector<int32_t> htc(nnz);
    KS_HIPCK(hipMemcpyAsync(htc.data(), m->tcol.p, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, s));
    KS_HIPCK(hipStreamSynchronize(s));
    build_panels(m, m->tpanels, htp.data(), htc.data(), m->tperm.p, m->flags);
  }
}

void ensure_fp32(ks_mat* m) {
  ks_ctx* c = m->ctx;
  const int64_t nnz = m->nnz;
  if (!m->val32.p) {
    m->val32.alloc(std::max<int64_t>(nnz, 1));
    k_convert<double, float><<<grid_map(nnz), kBlock, 0, c->stream>>>(nnz, m->val.p, m->val32.p, Gate{});
    KS_HIPCK(hipGetLastError());
  }
  if (m->has_t && !m->tval32.p) {
    m->tval32.alloc(std::max<int64_t>(nnz, 1));
    k_convert<double, float><<<grid_map(nnz), kBlock, 0, c->stream>>>(nnz, m->tval.p, m->tval32.p, Gate{});
    KS_HIPCK(hipGetLastError());
  }
  panels_fp32(m, m->panels);
  panels_fp32(m, m->tpanels);
}

// ------------------------------------------------------------- levels ---
LevelCfg outer_cfg(const ks_params* p, int64_t hist_cap) {
  LevelCfg c;
  c.maxloop = p->outer_maxloop; c.eps = p->outer_eps; c.restart = p->outer_restart;
  c.kskip = p->outer_kskip; c.fix = p->outer_fix; c.compat = p->ref_compat; c.hist_cap = hist_cap;
  return c;
}
LevelCfg inner_cfg(const ks_params* p) {
  LevelCfg c;
  c.maxloop = p->inner_maxloop; c.eps = p->inner_eps; c.restart = p->inner_restart;
  c.kskip = p->inner_kskip; c.fix = p->inner_fix; c.compat = p->ref_compat; c.hist_cap = 0;
  return c;
}

template <class T>
std::unique_ptr<Level> make_plain(Name nm, ks_ctx* c, const MatView<T>& A, const MatView<T>& AT,
                                  const LevelCfg& cfg) {
  switch (nm) {
    case CG: {
      auto L = std::make_unique<CgLevel<T>>();
      L->setup(c, A, cfg);
      return L;
    }
    case BICG: {
      auto L = std::make_unique<BicgLevel<T>>();
      L->setup(c, A, AT, cfg);
      return L;
    }
    case GMRES: {
      auto L = std::make_unique<GmresLevel<T>>();
      L->setup(c, A, cfg);
      return L;
    }
    case CR: {
      auto L = std::make_unique<CrLevel<T>>();
      L->setup(c, A, cfg);
      return L;
    }
    case GCR: {
      auto L = std::make_unique<GcrLevel<T>>();
      L->setup(c, A, cfg);
      return L;
    }
    default: break;
  }
  if constexpr (sizeof(T) == 8) {
    if (nm == KSKIPCG) {
      auto L = std::make_unique<KskipLevel<double, false>>();
      L->setup(c, A, AT, cfg);
      return L;
    }
    if (nm == KSKIPBICG) {
      auto L = std::make_unique<KskipLevel<double, true>>();
      L->setup(c, A, AT, cfg);
      return L;
    }
  }
  throw Error(KS_ERR_UNSUPPORTED, "solver not supported in this configuration (precision " +
                                      std::to_string(sizeof(T) * 8) + ")");
}

bool needs_t(Name nm) { return nm == BICG || nm == KSKIPBICG; }

// Build the level tree for (outer, inner).  trans: solve with A^T.
std::unique_ptr<Level> make_solver(ks_ctx* c, ks_mat* m, const ks_params* p, bool trans, int64_t hist_cap) {
  const Name on = name_of(p->outer), in = name_of(p->inner);
  KS_CHECK(on != NONE, KS_ERR_BAD_INPUT, "OuterSolverName can not found in List");
  KS_CHECK(on != KSKIPCR && on != VPBICG, KS_ERR_UNSUPPORTED, "Not implemented");
  const bool vp = on == VPCG || on == VPGMRES || on == VPGCR || on == VPCR;
  if (vp) KS_CHECK(in != NONE, KS_ERR_BAD_INPUT, "VP solver needs an inner solver (-s)");
  // A^T is needed for BiCG-family solves, and for any adjoint solve unless
  // the matrix is symmetric
  bool want_t = needs_t(on) || (vp && needs_t(in));
  if (trans && !(m->flags & KS_MAT_SPD)) want_t = true;
  KS_CHECK(!want_t || m->has_t, KS_ERR_BAD_INPUT, "matrix was created without KS_MAT_TRANSPOSE");
  const bool useT = trans && !(m->flags & KS_MAT_SPD);
  auto A64 = m->view<double>(useT);
  auto AT64 = m->has_t ? m->view<double>(!useT) : A64;
  const LevelCfg oc = outer_cfg(p, hist_cap);
  if (!vp) return make_plain<double>(on, c, A64, AT64, oc);
  KS_CHECK(on == VPCG || on == VPGMRES || on == VPGCR, KS_ERR_UNSUPPORTED,
           "VPCR is not implemented (the reference marks it broken, README.md:15,32)");
  const LevelCfg ic = inner_cfg(p);
  std::unique_ptr<Level> inner;
  if (p->inner_precision == KS_F32) {
    ensure_fp32(m);
    auto A32 = m->view<float>(useT);
    auto AT32 = m->has_t ? m->view<float>(!useT) : A32;
    inner = make_plain<float>(in, c, A32, AT32, ic);
  } else {
    KS_CHECK(p->inner_precision == KS_F64 || p->inner_precision == 0, KS_ERR_BAD_INPUT,
             "inner_precision must be 64 or 32");
    inner = make_plain<double>(in, c, A64, AT64, ic);
  }
  if (on == VPCG) {
    auto L = std::make_unique<VpcgLevel>();
    L->setup(c, A64, oc, std::move(inner));
    return L;
  }
  if (on == VPGCR) {
    auto L = std::make_unique<GcrLevel<double>>();
    L->setup(c, A64, oc, std::move(inner));
    return L;
  }
  auto L = std::make_unique<GmresLevel<double>>();
  L->setup(c, A64, oc, std::move(inner));
  return L;
}

// Enqueue a whole outer solve with batched host checks of the stop flag.
void run_outer(ks_ctx* c, Level& L, const double* b, double* x, int check_every) {
  volatile int* hf = c->hflags;
  hf[0] = 0;
  L.hctl.host_stop = c->hflags_dev;
  L.upload_ctl();
  L.enqueue_init(b, x);
  const int64_t U = L.max_units();
  int64_t B = check_every;
  if (B <= 0) {
    const double est = std::max(1e-6, L.unit_seconds_estimate());
    B = (int64_t)std::ceil(4e-4 / est);
    B = std::min<int64_t>(std::max<int64_t>(B, 1), 4096);
  }
  // the events of the last two batches (the pool is pre-filled by
  // ks_ctx_create, so a solve creates none)
  hipEvent_t q[2] = {nullptr, nullptr};
  int nq = 0;
  int64_t u = 0;
  bool stopped = false;
  while (u < U && !stopped) {
    const int64_t nb = std::min(B, U - u);
    for (int64_t i = 0; i < nb; ++i) L.enqueue_unit(u++);
    hipEvent_t ev = ev_get(c);
    KS_HIPCK(hipEventRecord(ev, c->stream));
    q[nq++] = ev;
    if (nq == 2) {
      KS_HIPCK(hipEventSynchronize(q[0]));
      c->evpool.push_back(q[0]);
      q[0] = q[1];
      nq = 1;
      if (hf[0]) stopped = true;
    }
  }
  L.enqueue_finish();
  KS_HIPCK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < nq; ++i) c->evpool.push_back(q[i]);
}

// ||b - A x||_2 on the device (FP64); the residual goes to rbuf (n doubles)
// when given, else to a temporary.
double resid_norm(ks_ctx* c, ks_mat* m, bool trans, const double* b, const double* x, double* rbuf = nullptr) {
  DBuf<double> r;
  if (rbuf)
    r.borrow(rbuf, m->n);
  else
    r.alloc(m->n);
  auto A = m->view<double>(trans && m->has_t);
  launch_spmv<double, 1>(c, A, x, EpiResid<double, FinStore>{b, r.p, FinStore{c->scal.p + 1, 1}}, Gate{}, 8.0 * m->n);
  double v = 0;
  KS_HIPCK(hipMemcpyAsync(&v, c->scal.p + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return v;
}
double norm_dev(ks_ctx* c, int64_t n, const double* x) {
  launch_map<1>(c, n, OpDot<double>{x, x, c->scal.p + 2, 1}, Gate{}, KS_K_REDUCE, 8.0 * n);
  double v = 0;
  KS_HIPCK(hipMemcpyAsync(&v, c->scal.p + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return v;
}

int64_t total_inner_steps(Level& L) {
  Level* in = nullptr;
  if (auto* g = dynamic_cast<GmresLevel<double>*>(&L)) in = g->inner.get();
  if (auto* v = dynamic_cast<VpcgLevel*>(&L)) in = v->inner.get();
  if (auto* q = dynamic_cast<GcrLevel<double>*>(&L)) in = q->inner.get();
  if (!in) return 0;
  Ctl h;
  KS_HIPCK(hipMemcpy(&h, in->dctl.p, sizeof(Ctl), hipMemcpyDeviceToHost));
  return h.tsteps;
}

// Run a constructed solve level and fill the result.  prof_mask selects the
// kernel classes timed with events; rbuf (optional, n doubles of device
// memory the solve no longer needs at the end) holds the residuals of the
// true-residual checks.
void run_level(ks_ctx* c, ks_mat* m, Level& Lv, const ks_params* p, unsigned prof_mask, bool trans,
               const double* b, double* x, double* hist, int64_t hist_cap, ks_result* res, double* x0norm_out,
               double* rbuf) {
  Level* L = &Lv;
  L->x0given = (p->flags & KS_X0_GIVEN) != 0;
  double r0 = 0;
  if (L->x0given) r0 = resid_norm(c, m, trans, b, x, rbuf);
  k_fill<unsigned long long><<<1, kBlock, 0, c->stream>>>(2 * KS_K_NCLASS, c->stats.p, 0ull);
  KS_HIPCK(hipGetLastError());
  c->prof.mask = prof_mask;
  KS_HIPCK(hipStreamSynchronize(c->stream));
  const auto t0 = std::chrono::steady_clock::now();
  run_outer(c, *L, b, x, p->check_every);
  const auto t1 = std::chrono::steady_clock::now();
  Ctl h;
  KS_HIPCK(hipMemcpy(&h, L->dctl.p, sizeof(Ctl), hipMemcpyDeviceToHost));
  unsigned long long st[2 * KS_K_NCLASS];
  KS_HIPCK(hipMemcpy(st, c->stats.p, sizeof(st), hipMemcpyDeviceToHost));
  double ms[KS_K_NCLASS] = {0};
  c->prof.collect(ms);
  c->prof.mask = 0;
  if (res) {
    std::memset(res, 0, sizeof(*res));
    res->status = h.stop ? h.status : KS_NOT_CONVERGED;
    res->iters = L->reported_iters(h);
    res->steps = h.tsteps;
    res->inner_steps = total_inner_steps(*L);
    res->spmv = (int64_t)(st[KS_K_SPMV64] + st[KS_K_SPMV32] + 2 * st[KS_K_BASIS]);
    res->relres = h.err;
    res->seconds = std::chrono::duration<double>(t1 - t0).count();
    double tot = 0;
    for (int k = 0; k < KS_K_NCLASS; ++k) {
      res->prof_ms[k] = ms[k];
      res->prof_launches[k] = (int64_t)st[k];
      res->prof_bytes[k] = (double)st[KS_K_NCLASS + k];
      tot += (double)st[KS_K_NCLASS + k];
    }
    res->bytes = tot;
    const double rn = resid_norm(c, m, trans, b, x, rbuf);
    const double bn = norm_dev(c, m->n, b);
    res->relres_true = bn > 0 ? rn / bn : rn;
    const double den = L->x0given ? r0 : bn;
    res->log10_check = std::log10(rn / den);
    if (hist && hist_cap > 0) {
      const int64_t nh = std::min<int64_t>(h.nhist, hist_cap);
      if (nh > 0) KS_HIPCK(hipMemcpy(hist, L->hist.p, sizeof(double) * 2 * nh, hipMemcpyDeviceToHost));
      res->n_hist = h.nhist;
    }
  }
  if (x0norm_out) *x0norm_out = r0;
}

void solve_impl(ks_ctx* c, ks_mat* m, const ks_params* p, bool trans, const double* b, double* x, double* hist,
                int64_t hist_cap, ks_result* res, double* x0norm_out) {
  auto L = make_solver(c, m, p, trans, hist ? hist_cap : 0);
  run_level(c, m, *L, p, p->profile > 0 ? (unsigned)p->profile : 0u, trans, b, x, hist, hist_cap, res, x0norm_out,
            nullptr);
}

// ------------------------------------------------ no-allocation CG solve ---
// ks_solve_noalloc keeps every piece of per-call device state of a plain FP64
// CG solve (r, p, q, the small arrays and the control block) in `scratch`, a
// file-scope device array of kNoAllocMaxRows-row capacity (3 * 2^20 doubles
// of vectors + the control state, about 24 MB held by each GPU the library
// runs on until the process ends).  In host code `scratch` names a host-side
// stand-in, never the device memory: its device address on each GPU comes
// from hipGetSymbolAddress, taken once per device with that device current
// (scratch_address, from ks_ctx_create), and only that address is passed to
// kernels.  Calls on one GPU share the array: a per-device lock serialises
// them within the process, and each call clears what must start at zero.
constexpr int64_t kNoAllocMaxRows = int64_t(1) << 20;
constexpr int kScratchMaxDevices = 64;
__device__ __align__(256) double scratch[3 * kNoAllocMaxRows + 1024];

struct ScratchTable {
  std::mutex mu;                       // guards addr
  double* addr[kScratchMaxDevices] = {};
  std::mutex call[kScratchMaxDevices];  // one ks_solve_noalloc at a time per device
};
ScratchTable& scratch_table() {
  static ScratchTable t;
  return t;
}

// Device address of `scratch` on `device` (which must be current); nullptr
// for device numbers beyond the table.
double* scratch_address(int device) {
  if (device < 0 || device >= kScratchMaxDevices) return nullptr;
  ScratchTable& t = scratch_table();
  std::lock_guard<std::mutex> lk(t.mu);
  if (!t.addr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw Error(KS_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    t.addr[device] = (double*)ptr;
  }
  return t.addr[device];
}

}  // namespace

#define KS_API_TRY try {
#define KS_API_CATCH(ctxp)                                           \
  }                                                                  \
  catch (const ks::Error& e) {                                       \
    if (ctxp) (ctxp)->err = e.what();                                \
    return (ks_status)e.status;                                      \
  }                                                                  \
  catch (const std::exception& e) {                                  \
    if (ctxp) (ctxp)->err = e.what();                                \
    return KS_ERR_BAD_INPUT;                                         \
  }

extern "C" {

ks_status ks_ctx_create(int32_t device, ks_ctx** out) {
  if (!out) return KS_ERR_BAD_INPUT;
  *out = nullptr;
  ks_ctx* c = new ks_ctx();
  try {
    int nd = 0;
    KS_HIPCK(hipGetDeviceCount(&nd));
    KS_CHECK(device >= 0 && device < nd, KS_ERR_BAD_INPUT, "no HIP device " + std::to_string(device));
    c->device = device;
    KS_HIPCK(hipSetDevice(device));
    KS_HIPCK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->part.alloc((size_t)kMaxGrid * kMaxNV);
    c->part2.alloc((size_t)kMaxGrid);
    KS_HIPCK(hipMemsetAsync(c->part2.p, 0, sizeof(double) * kMaxGrid, c->stream));
    c->stats.alloc(2 * KS_K_NCLASS);
    c->scal.alloc(1024);
    KS_HIPCK(hipHostMalloc((void**)&c->hflags, 64 * sizeof(int), hipHostMallocMapped));
    std::memset(c->hflags, 0, 64 * sizeof(int));
    KS_HIPCK(hipHostGetDevicePointer((void**)&c->hflags_dev, c->hflags, 0));
    hipDeviceProp_t prop;
    KS_HIPCK(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    // the batch events of run_outer (it holds at most three at a time), so
    // that a solve creates none
    c->evpool.reserve(16);
    for (int i = 0; i < 4; ++i) {
      hipEvent_t e;
      KS_HIPCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      c->evpool.push_back(e);
    }
    c->scratch = scratch_address(device);
    KS_HIPCK(hipStreamSynchronize(c->stream));
  } catch (const ks::Error& e) {
    int st = e.status;
    ks_ctx_destroy(c);
    return (ks_status)st;
  }
  *out = c;
  return KS_OK;
}

void ks_ctx_destroy(ks_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->prof.destroy();
  for (auto e : c->evpool) (void)hipEventDestroy(e);
  c->part.release();
  c->stats.release();
  c->scal.release();
  if (c->hflags) (void)hipHostFree(c->hflags);
  if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* ks_ctx_last_error(const ks_ctx* c) { return c ? c->err.c_str() : "null context"; }
void* ks_ctx_stream(ks_ctx* c) { return c ? (void*)c->stream : nullptr; }
ks_status ks_ctx_synchronize(ks_ctx* c) {
  if (!c) return KS_ERR_BAD_INPUT;
  KS_API_TRY
  set_device(c);
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_mat_create_csr(ks_ctx* c, int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col,
                            const double* val, uint32_t flags, ks_mat** out) {
  if (!c || !out) return KS_ERR_BAD_INPUT;
  *out = nullptr;
  auto m = std::make_unique<ks_mat>();
  KS_API_TRY
  set_device(c);
  KS_CHECK(n > 0 && n < (int64_t)INT32_MAX, KS_ERR_BAD_INPUT, "n must be in [1, 2^31)");
  KS_CHECK(nnz >= 0 && nnz < (int64_t)INT32_MAX, KS_ERR_BAD_INPUT, "nnz must be in [0, 2^31)");
  KS_CHECK(ptr && (nnz == 0 || (col && val)), KS_ERR_BAD_INPUT, "null CSR array");
  m->ctx = c; m->n = n; m->nnz = nnz; m->flags = flags;
  if (flags & KS_MAT_BORROW) {
    KS_CHECK(flags & KS_PTR_DEVICE, KS_ERR_BAD_INPUT, "KS_MAT_BORROW needs device arrays (KS_PTR_DEVICE)");
    KS_CHECK(nnz > 0, KS_ERR_BAD_INPUT, "KS_MAT_BORROW needs nnz > 0");
    // the SpMV reads values and column indices in aligned pairs
    KS_CHECK((uintptr_t)col % 16 == 0 && (uintptr_t)val % 16 == 0, KS_ERR_BAD_INPUT,
             "KS_MAT_BORROW needs 16-byte aligned col and val arrays");
    m->ptr.borrow(const_cast<int32_t*>(ptr), n + 1);
    m->col.borrow(const_cast<int32_t*>(col), nnz);
    m->val.borrow(const_cast<double*>(val), nnz);
  } else {
    m->ptr.alloc(n + 1);
    m->col.alloc(std::max<int64_t>(nnz, 1));
    m->val.alloc(std::max<int64_t>(nnz, 1));
    const hipMemcpyKind k = (flags & KS_PTR_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    KS_HIPCK(hipMemcpyAsync(m->ptr.p, ptr, sizeof(int32_t) * (n + 1), k, c->stream));
    if (nnz > 0) {
      KS_HIPCK(hipMemcpyAsync(m->col.p, col, sizeof(int32_t) * nnz, k, c->stream));
      KS_HIPCK(hipMemcpyAsync(m->val.p, val, sizeof(double) * nnz, k, c->stream));
    }
  }
  // validate on the host (row pointers, column range)
  std::vector<int32_t> hp(n + 1), hc;
  KS_HIPCK(hipMemcpyAsync(hp.data(), m->ptr.p, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
  if (nnz > 0) {
    hc.resize(nnz);
    KS_HIPCK(hipMemcpyAsync(hc.data(), m->col.p, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, c->stream));
  }
  KS_HIPCK(hipStreamSynchronize(c->stream));
  KS_CHECK(hp[0] == 0 && hp[n] == nnz, KS_ERR_BAD_INPUT, "row pointers must start at 0 and end at nnz");
  int64_t mx = 0;
  for (int64_t i = 0; i < n; ++i) {
    KS_CHECK(hp[i + 1] >= hp[i], KS_ERR_BAD_INPUT, "row pointers must be non-decreasing");
    mx = std::max<int64_t>(mx, hp[i + 1] - hp[i]);
  }
  for (int64_t j = 0; j < nnz; ++j)
    KS_CHECK(hc[j] >= 0 && hc[j] < n, KS_ERR_BAD_INPUT, "column index out of range at entry " + std::to_string(j));
  m->max_row = mx;
  KS_CHECK(!((flags & KS_MAT_COL32) && (flags & KS_MAT_COL_DELTA)), KS_ERR_BAD_INPUT,
           "KS_MAT_COL32 and KS_MAT_COL_DELTA exclude each other");
  const bool off16 = !(flags & KS_MAT_COL32) && nnz > 0 && offsets_fit16(n, hp.data(), hc.data());
  make_plan(c, m->plan, hp.data(), n, nnz, (flags & KS_MAT_VECTOR) != 0, tile_choice(flags),
              merge_choice(flags, n, nnz, mx));
  if (nnz > 0 && panel_choice(flags, n, hp.data(), hc.data()))
    build_panels(m.get(), m->panels, hp.data(), hc.data(), nullptr, flags);
  if (flags & KS_MAT_TRANSPOSE) build_transpose(m.get());
  if (!(flags & KS_MAT_COL32) && nnz > 0) build_index16(m.get(), off16);
  if (flags & KS_MAT_FP32_COPY) ensure_fp32(m.get());
  KS_HIPCK(hipStreamSynchronize(c->stream));
  *out = m.release();
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_mat_update_values(ks_mat* m, const double* val, uint32_t flags) {
  if (!m) return KS_ERR_BAD_INPUT;
  ks_ctx* c = m->ctx;
  KS_API_TRY
  set_device(c);
  if (m->nnz == 0) return KS_OK;
  KS_CHECK(val, KS_ERR_BAD_INPUT, "null values");
  // a borrowed matrix updated in place passes its own value array: only the
  // derived copies (A^T values, FP32 values) are refreshed
  if (val != m->val.p) {
    const hipMemcpyKind k = (flags & KS_PTR_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    KS_HIPCK(hipMemcpyAsync(m->val.p, val, sizeof(double) * m->nnz, k, c->stream));
  }
  if (m->has_t) {
    k_gather<double, double><<<grid_map(m->nnz), kBlock, 0, c->stream>>>(m->nnz, m->tperm.p, m->val.p, m->tval.p);
    KS_HIPCK(hipGetLastError());
  }
  refresh_panels(m, m->panels);
  refresh_panels(m, m->tpanels);
  if (m->val32.p) {
    k_convert<double, float><<<grid_map(m->nnz), kBlock, 0, c->stream>>>(m->nnz, m->val.p, m->val32.p, Gate{});
    KS_HIPCK(hipGetLastError());
  }
  if (m->tval32.p) {
    k_convert<double, float><<<grid_map(m->nnz), kBlock, 0, c->stream>>>(m->nnz, m->tval.p, m->tval32.p, Gate{});
    KS_HIPCK(hipGetLastError());
  }
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_mat_get_transpose(ks_mat* m, int32_t* tptr, int32_t* tcol, double* tval, uint32_t flags) {
  if (!m) return KS_ERR_BAD_INPUT;
  ks_ctx* c = m->ctx;
  KS_API_TRY
  set_device(c);
  KS_CHECK(m->has_t, KS_ERR_BAD_INPUT, "matrix was created without KS_MAT_TRANSPOSE");
  const hipMemcpyKind k = (flags & KS_PTR_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  KS_HIPCK(hipMemcpyAsync(tptr, m->tptr.p, sizeof(int32_t) * (m->n + 1), k, c->stream));
  if (m->nnz > 0) {
    KS_HIPCK(hipMemcpyAsync(tcol, m->tcol.p, sizeof(int32_t) * m->nnz, k, c->stream));
    KS_HIPCK(hipMemcpyAsync(tval, m->tval.p, sizeof(double) * m->nnz, k, c->stream));
  }
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

void ks_mat_destroy(ks_mat* m) {
  if (!m) return;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  delete m;
}

ks_status ks_mat_info(const ks_mat* m, int64_t* n, int64_t* nnz, int32_t* kind, int32_t* lanes) {
  if (!m) return KS_ERR_BAD_INPUT;
  if (n) *n = m->n;
  if (nnz) *nnz = m->nnz;
  if (kind) *kind = m->panels.np ? 3 : m->plan.kind;
  if (lanes) *lanes = m->plan.L;
  return KS_OK;
}

ks_status ks_mat_index_bytes(const ks_mat* m, int32_t trans, int32_t precision, int32_t* bytes) {
  if (!m || !bytes || (precision != KS_F64 && precision != KS_F32)) return KS_ERR_BAD_INPUT;
  if (trans && !m->has_t) return KS_ERR_BAD_INPUT;
  *bytes = precision == KS_F64 ? m->view<double>(trans != 0).ib : m->view<float>(trans != 0).ib;
  return KS_OK;
}

ks_status ks_spmv(ks_mat* m, int32_t trans, int32_t precision, const void* x, void* y, uint32_t flags) {
  if (!m) return KS_ERR_BAD_INPUT;
  ks_ctx* c = m->ctx;
  KS_API_TRY
  set_device(c);
  KS_CHECK(!trans || m->has_t, KS_ERR_BAD_INPUT, "matrix was created without KS_MAT_TRANSPOSE");
  if (precision == KS_F32) {
    ensure_fp32(m);
    In<float> xi(c, (const float*)x, m->n, flags);
    Out<float> yo(c, (float*)y, m->n, flags);
    launch_spmv<float, 0>(c, m->view<float>(trans != 0), xi.p, EpiStream<float>{yo.p}, Gate{});
    yo.copy_back(c);
  } else {
    KS_CHECK(precision == KS_F64, KS_ERR_BAD_INPUT, "precision must be 64 or 32");
    In<double> xi(c, (const double*)x, m->n, flags);
    Out<double> yo(c, (double*)y, m->n, flags);
    launch_spmv<double, 0>(c, m->view<double>(trans != 0), xi.p, EpiStream<double>{yo.p}, Gate{});
    yo.copy_back(c);
  }
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

// A pair of timing events, destroyed on every exit path.
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  EventPair() {
    KS_HIPCK(hipEventCreate(&a));
    KS_HIPCK(hipEventCreate(&b));
  }
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  EventPair(const EventPair&) = delete;
  EventPair& operator=(const EventPair&) = delete;
};

ks_status ks_spmv_bench(ks_mat* m, int32_t trans, int32_t precision, const void* x, void* y, int32_t reps,
                        double* ms_per_launch) {
  if (!m || reps <= 0) return KS_ERR_BAD_INPUT;
  ks_ctx* c = m->ctx;
  KS_API_TRY
  set_device(c);
  KS_CHECK(!trans || m->has_t, KS_ERR_BAD_INPUT, "matrix was created without KS_MAT_TRANSPOSE");
  if (precision == KS_F32) ensure_fp32(m);
  EventPair ev;
  KS_HIPCK(hipEventRecord(ev.a, c->stream));
  for (int r = 0; r < reps; ++r) {
    if (precision == KS_F32)
      launch_spmv<float, 0>(c, m->view<float>(trans != 0), (const float*)x, EpiStream<float>{(float*)y}, Gate{});
    else
      launch_spmv<double, 0>(c, m->view<double>(trans != 0), (const double*)x, EpiStream<double>{(double*)y}, Gate{});
  }
  KS_HIPCK(hipEventRecord(ev.b, c->stream));
  KS_HIPCK(hipEventSynchronize(ev.b));
  float t = 0;
  KS_HIPCK(hipEventElapsedTime(&t, ev.a, ev.b));
  if (ms_per_launch) *ms_per_launch = t / reps;
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_bench_stream(ks_ctx* c, int32_t kind, int64_t bytes, int32_t reps, double* gbps) {
  if (!c || !gbps || kind < 0 || kind > 1 || reps <= 0 || bytes < 16 * 1024 || bytes % 16 != 0 ||
      bytes > (int64_t)16 << 30)
    return KS_ERR_BAD_INPUT;
  KS_API_TRY
  set_device(c);
  const int64_t n16 = bytes / 16;
  static constexpr int kGrids[] = {1024, 2048, 4096, 8192};
  DBuf<StreamChunk> x(n16), y(kind == 1 ? n16 : 0);
  DBuf<float> part((size_t)kGrids[3] * (kBlock / 64));
  k_stream_fill<<<grid_map(n16), kBlock, 0, c->stream>>>(n16, x.p, 1.0f);
  KS_HIPCK(hipGetLastError());
  EventPair ev;
  double best = 0.0;
  for (const int G : kGrids) {
    auto go = [&] {
      if (kind == 0) k_stream_read<<<G, kBlock, 0, c->stream>>>(n16, x.p, part.p);
      else k_stream_copy<<<G, kBlock, 0, c->stream>>>(n16, x.p, y.p);
    };
    for (int r = 0; r < 2; ++r) go();
    KS_HIPCK(hipEventRecord(ev.a, c->stream));
    for (int r = 0; r < reps; ++r) go();
    KS_HIPCK(hipEventRecord(ev.b, c->stream));
    KS_HIPCK(hipEventSynchronize(ev.b));
    KS_HIPCK(hipGetLastError());
    float t = 0;
    KS_HIPCK(hipEventElapsedTime(&t, ev.a, ev.b));
    const double moved = (double)bytes * (kind == 1 ? 2 : 1) * reps;
    if (t > 0) best = std::max(best, moved / (t * 1e-3) / 1e9);
  }
  *gbps = best;
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_dot(ks_ctx* c, int64_t n, int32_t precision, const void* x, const void* y, double* out,
                 uint32_t flags) {
  if (!c || !out || n < 0) return KS_ERR_BAD_INPUT;
  KS_API_TRY
  set_device(c);
  if (precision == KS_F32) {
    In<float> xi(c, (const float*)x, n, flags), yi(c, (const float*)y, n, flags);
    launch_map<1>(c, n, OpDot<float>{xi.p, yi.p, c->scal.p, 0}, Gate{}, KS_K_REDUCE, 8.0 * n);
  } else {
    In<double> xi(c, (const double*)x, n, flags), yi(c, (const double*)y, n, flags);
    launch_map<1>(c, n, OpDot<double>{xi.p, yi.p, c->scal.p, 0}, Gate{}, KS_K_REDUCE, 16.0 * n);
  }
  KS_HIPCK(hipMemcpyAsync(out, c->scal.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KS_HIPCK(hipStreamSynchronize(c->stream));
  if (n == 0) *out = 0.0;
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_norm2(ks_ctx* c, int64_t n, int32_t precision, const void* x, double* out, uint32_t flags) {
  if (!c || !out || n < 0) return KS_ERR_BAD_INPUT;
  KS_API_TRY
  set_device(c);
  if (precision == KS_F32) {
    In<float> xi(c, (const float*)x, n, flags);
    launch_map<1>(c, n, OpDot<float>{xi.p, xi.p, c->scal.p, 1}, Gate{}, KS_K_REDUCE, 4.0 * n);
  } else {
    In<double> xi(c, (const double*)x, n, flags);
    launch_map<1>(c, n, OpDot<double>{xi.p, xi.p, c->scal.p, 1}, Gate{}, KS_K_REDUCE, 8.0 * n);
  }
  KS_HIPCK(hipMemcpyAsync(out, c->scal.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KS_HIPCK(hipStreamSynchronize(c->stream));
  if (n == 0) *out = 0.0;
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_axpy(ks_ctx* c, int64_t n, double a, const double* x, const double* y, double* out,
                  uint32_t flags) {
  if (!c || n < 0) return KS_ERR_BAD_INPUT;
  KS_API_TRY
  set_device(c);
  In<double> xi(c, x, n, flags), yi(c, y, n, flags);
  Out<double> o(c, out, n, flags);
  launch_map<0>(c, n, OpAxpyApi{a, xi.p, yi.p, o.p}, Gate{}, KS_K_AXPY, 24.0 * n);
  o.copy_back(c);
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_cgs(ks_ctx* c, int64_t n, int32_t k, const double* V, double* w, double* h, uint32_t flags) {
  if (!c || n <= 0 || k < 0 || !h) return KS_ERR_BAD_INPUT;
  KS_API_TRY
  set_device(c);
  In<double> Vi(c, V, n * (int64_t)(k + 1), flags);
  Out<double> wo(c, w, n, flags, true);
  DBuf<double> hd(k + 2);
  launch_proj<double>(c, n, Vi.p, n, k + 1, wo.p, hd.p, 1, Gate{});
  launch_cgs_update<double>(c, n, Vi.p, n, k + 1, wo.p, wo.p, hd.p, 1, FinStore{hd.p + k + 1, 1}, Gate{});
  wo.copy_back(c);
  KS_HIPCK(hipMemcpyAsync(h, hd.p, sizeof(double) * (k + 2), hipMemcpyDeviceToHost, c->stream));
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_kskip_basis(ks_mat* m, int32_t k, const double* r, const double* p, double* Ar, double* Ap,
                         uint32_t flags) {
  if (!m || k < 0) return KS_ERR_BAD_INPUT;
  ks_ctx* c = m->ctx;
  KS_API_TRY
  set_device(c);
  const int64_t n = m->n;
  In<double> ri(c, r, n, flags), pi(c, p, n, flags);
  Out<double> Aro(c, Ar, n * (2 * k + 1), flags), Apo(c, Ap, n * (2 * k + 2), flags);
  auto A = m->view<double>(false);
  for (int ii = 0; ii < 2 * k + 2; ++ii) {
    const double* sr = ii == 0 ? ri.p : Aro.p + (int64_t)(ii - 1) * n;
    const double* sp = ii == 0 ? pi.p : Apo.p + (int64_t)(ii - 1) * n;
    if (ii < 2 * k + 1)
      launch_spmv2<double>(c, A, sr, sp, Aro.p + (int64_t)ii * n, Apo.p + (int64_t)ii * n, Gate{});
    else
      launch_spmv<double, 0>(c, A, sp, EpiStore<double>{Apo.p + (int64_t)ii * n}, Gate{});
  }
  Aro.copy_back(c);
  Apo.copy_back(c);
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_kskip_gram(ks_ctx* c, int64_t n, int32_t k, const double* r, const double* p, const double* Ar,
                        const double* Ap, double* delta, double* eta, double* zeta, uint32_t flags) {
  if (!c || n <= 0 || k < 1 || k > 7) return k > 7 ? KS_ERR_UNSUPPORTED : KS_ERR_BAD_INPUT;
  KS_API_TRY
  set_device(c);
  In<double> ri(c, r, n, flags), pi(c, p, n, flags), Ari(c, Ar, n * (2 * k + 1), flags),
      Api(c, Ap, n * (2 * k + 2), flags);
  DBuf<Ctl> dc(1);
  DBuf<double> arr(3 * (2 * k + 3));
  Ctl h{};
  h.delta = arr.p;
  h.eta = arr.p + (2 * k + 3);
  h.zeta = arr.p + 2 * (2 * k + 3);
  KS_HIPCK(hipMemcpyAsync(dc.p, &h, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
#define KS_GRAM(KK)                                                                                  \
  case KK: {                                                                                         \
    OpGram<KK, false> op{n, ri.p, pi.p, nullptr, nullptr, Ari.p, Api.p, dc.p};                       \
    launch_map<OpGram<KK, false>::NV>(c, n, op, Gate{}, KS_K_BASIS, 8.0 * n * (4 * KK + 5));         \
  } break;
  switch (k) {
    KS_GRAM(1) KS_GRAM(2) KS_GRAM(3) KS_GRAM(4) KS_GRAM(5) KS_GRAM(6) KS_GRAM(7)
  }
#undef KS_GRAM
  std::vector<double> out(3 * (2 * k + 3));
  KS_HIPCK(hipMemcpyAsync(out.data(), arr.p, sizeof(double) * out.size(), hipMemcpyDeviceToHost, c->stream));
  KS_HIPCK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < 2 * k; ++i) delta[i] = out[i];
  for (int i = 0; i < 2 * k + 1; ++i) eta[i] = out[(2 * k + 3) + i];
  for (int i = 0; i < 2 * k + 2; ++i) zeta[i] = out[2 * (2 * k + 3) + i];
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_sddmm(ks_mat* m, double alpha, const double* u, const double* v, double* out_vals, uint32_t flags) {
  if (!m) return KS_ERR_BAD_INPUT;
  ks_ctx* c = m->ctx;
  KS_API_TRY
  set_device(c);
  const int64_t n = m->n;
  In<double> ui(c, u, n, flags), vi(c, v, n, flags);
  Out<double> o(c, out_vals, m->nnz, flags);
  const int L = m->plan.L;
  const int64_t nrb = (n + kBlock / L - 1) / (kBlock / L);
  const double bytes = 12.0 * m->nnz + 4.0 * (n + 1) + 16.0 * n;
  const int G = grid_rows(nrb);
  switch (L) {
    case 2: k_sddmm<2><<<G, kBlock, 0, c->stream>>>(n, m->ptr.p, m->col.p, alpha, ui.p, vi.p, o.p, stat(c, KS_K_OTHER, bytes)); break;
    case 4: k_sddmm<4><<<G, kBlock, 0, c->stream>>>(n, m->ptr.p, m->col.p, alpha, ui.p, vi.p, o.p, stat(c, KS_K_OTHER, bytes)); break;
    case 8: k_sddmm<8><<<G, kBlock, 0, c->stream>>>(n, m->ptr.p, m->col.p, alpha, ui.p, vi.p, o.p, stat(c, KS_K_OTHER, bytes)); break;
    case 16: k_sddmm<16><<<G, kBlock, 0, c->stream>>>(n, m->ptr.p, m->col.p, alpha, ui.p, vi.p, o.p, stat(c, KS_K_OTHER, bytes)); break;
    case 32: k_sddmm<32><<<G, kBlock, 0, c->stream>>>(n, m->ptr.p, m->col.p, alpha, ui.p, vi.p, o.p, stat(c, KS_K_OTHER, bytes)); break;
    default: k_sddmm<64><<<G, kBlock, 0, c->stream>>>(n, m->ptr.p, m->col.p, alpha, ui.p, vi.p, o.p, stat(c, KS_K_OTHER, bytes)); break;
  }
  KS_HIPCK(hipGetLastError());
  o.copy_back(c);
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_solve(ks_ctx* c, ks_mat* m, const ks_params* p, const double* b, double* x, double* hist,
                   int64_t hist_cap, ks_result* res) {
  if (!c || !m || !p || !b || !x) return KS_ERR_BAD_INPUT;
  KS_API_TRY
  set_device(c);
  KS_CHECK(m->ctx == c, KS_ERR_BAD_INPUT, "matrix belongs to another context");
  In<double> bi(c, b, m->n, p->flags);
  Out<double> xo(c, x, m->n, p->flags, (p->flags & KS_X0_GIVEN) != 0);
  solve_impl(c, m, p, false, bi.p, xo.p, hist, hist_cap, res, nullptr);
  xo.copy_back(c);
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_solve_noalloc(ks_ctx* c, ks_mat* m, const ks_params* p, const double* b, double* x,
                           ks_result* res) {
  if (!c || !m || !p || !b || !x) return KS_ERR_BAD_INPUT;
  KS_API_TRY
  set_device(c);
  KS_CHECK(m->ctx == c, KS_ERR_BAD_INPUT, "matrix belongs to another context");
  KS_CHECK(p->flags & KS_PTR_DEVICE, KS_ERR_BAD_INPUT,
           "ks_solve_noalloc takes device pointers only (KS_PTR_DEVICE): host arrays would need staging buffers");
  KS_CHECK(name_of(p->outer) == CG && name_of(p->inner) == NONE, KS_ERR_UNSUPPORTED,
           "ks_solve_noalloc runs plain CG only (outer \"cg\", inner \"NO\"); use ks_solve");
  KS_CHECK(m->n <= kNoAllocMaxRows, KS_ERR_BAD_INPUT,
           "ks_solve_noalloc: n = " + std::to_string(m->n) + " exceeds its maximum of " +
               std::to_string(kNoAllocMaxRows) + " rows (the size of its fixed scratch memory); use ks_solve");
  KS_CHECK(c->scratch != nullptr, KS_ERR_UNSUPPORTED, "no scratch memory on device " + std::to_string(c->device));
  static_assert(CgLevel<double>::scratch_doubles(kNoAllocMaxRows) <= 3 * kNoAllocMaxRows + 1024,
                "scratch holds three padded vectors and the control state at the maximum size");
  std::lock_guard<std::mutex> lk(scratch_table().call[c->device]);
  CgLevel<double> L;
  L.setup(c, m->view<double>(false), outer_cfg(p, 0), c->scratch);
  // q is free once the solve has ended: it holds the final residual check
  run_level(c, m, L, p, 0u, false, b, x, nullptr, 0, res, nullptr, L.q.p);
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

ks_status ks_adjoint(ks_ctx* c, ks_mat* m, const ks_params* p, const double* x, const double* dLdx,
                     double* grad_b, double* grad_vals, ks_result* res) {
  if (!c || !m || !p || !x || !dLdx || !grad_b) return KS_ERR_BAD_INPUT;
  KS_API_TRY
  set_device(c);
  KS_CHECK(m->ctx == c, KS_ERR_BAD_INPUT, "matrix belongs to another context");
  ks_params q = *p;
  q.flags &= ~KS_X0_GIVEN;
  In<double> xi(c, x, m->n, p->flags), gi(c, dLdx, m->n, p->flags);
  Out<double> gb(c, grad_b, m->n, p->flags);
  // A^T lambda = dL/dx
  solve_impl(c, m, &q, true, gi.p, gb.p, nullptr, 0, res, nullptr);
  if (grad_vals) {
    Out<double> gv(c, grad_vals, m->nnz, p->flags);
    KS_HIPCK(hipStreamSynchronize(c->stream));
    ks_status st = ks_sddmm(m, -1.0, gb.p, xi.p, gv.p, KS_PTR_DEVICE);
    if (st != KS_OK) return st;
    gv.copy_back(c);
  }
  gb.copy_back(c);
  KS_HIPCK(hipStreamSynchronize(c->stream));
  return KS_OK;
  KS_API_CATCH(c)
}

}  // extern "C"
