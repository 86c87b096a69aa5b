// ks_dev.h - host-side infrastructure of libksolve: device buffers, the
// context (stream, reduction scratch, statistics), the matrix (CSR in HBM,
// FP32 value copy, device-built transpose) and the kernel launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ksolve.h"
#include "ks_kernels.h"

namespace ks {

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define KS_HIPCK(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::ks::Error(KS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) +     \
                                        " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define KS_CHECK(cond, status, msg)                     \
  do {                                                  \
    if (!(cond)) throw ::ks::Error((status), (msg));    \
  } while (0)

// Owning device array.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  bool own = true;  // false: caller's memory (borrow), never freed here
  DBuf() = default;
  explicit DBuf(size_t m) { alloc(m); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), own(o.own) { o.p = nullptr; o.n = 0; o.own = true; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; own = o.own; o.p = nullptr; o.n = 0; o.own = true; }
    return *this;
  }
  void borrow(T* q, size_t m) {
    release();
    p = q;
    n = m;
    own = false;
  }
  ~DBuf() { release(); }
  void alloc(size_t m) {
    release();
    if (m == 0) return;
    hipError_t e = hipMalloc((void**)&p, m * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw Error(KS_ERR_NO_MEMORY, "hipMalloc of " + std::to_string(m * sizeof(T)) + " bytes failed: " +
                                        hipGetErrorString(e));
    }
    n = m;
  }
  void release() {
    if (p && own) (void)hipFree(p);
    p = nullptr;
    n = 0;
    own = true;
  }
};

// Per-class hipEvent timing of launches (profile mode).
struct Prof {
  unsigned mask = 0;
  struct Pair { hipEvent_t a, b; int cls; };
  std::vector<hipEvent_t> pool;
  std::vector<Pair> pend;
  size_t used = 0;
  int cur = -1;
  hipEvent_t get() {
    if (used == pool.size()) {
      hipEvent_t e;
      KS_HIPCK(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  void begin(hipStream_t s, int cls) {
    cur = -1;
    if (!(mask & (1u << cls))) return;
    Pair pr{get(), get(), cls};
    KS_HIPCK(hipEventRecord(pr.a, s));
    pend.push_back(pr);
    cur = cls;
  }
  void end(hipStream_t s) {
    if (cur < 0) return;
    KS_HIPCK(hipEventRecord(pend.back().b, s));
    cur = -1;
  }
  // an event pair for one launch of class cls (false: cls not profiled)
  bool pair(int cls, hipEvent_t& a, hipEvent_t& b) {
    if (!(mask & (1u << cls))) return false;
    Pair pr{get(), get(), cls};
    pend.push_back(pr);
    a = pr.a;
    b = pr.b;
    return true;
  }
  // after the stream is synchronised
  void collect(double* ms) {
    for (auto& pr : pend) {
      float t = 0.f;
      KS_HIPCK(hipEventElapsedTime(&t, pr.a, pr.b));
      ms[pr.cls] += t;
    }
    pend.clear();
    used = 0;
  }
  void destroy() {
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
    pend.clear();
    used = 0;
  }
};

}  // namespace ks

struct ks_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  std::string err;
  ks::DBuf<double> part;                 // kMaxGrid * kMaxNV reduction partials
  ks::DBuf<double> part2;                // kMaxGrid partials of a norm whose finisher is deferred
  ks::DBuf<unsigned long long> stats;    // launches[8], bytes[8]
  ks::DBuf<double> scal;                 // scalar outputs of the BLAS entry points
  int* hflags = nullptr;                 // host-mapped flags (host view)
  int* hflags_dev = nullptr;             // same memory, device view
  int num_cus = 256;
  ks::Prof prof;
  std::vector<hipEvent_t> evpool;
  double* scratch = nullptr;             // device address of the no-allocation solve's scratch array
  // resident workgroups per CU of each kernel (and dynamic LDS size) launched
  // on this context's device
  std::map<std::pair<const void*, size_t>, int> occ;
};

namespace ks {

// Tile (or merge-path segment) starts of one CSR-stream layout.
struct TileSet {
  int64_t ntiles = 0;
  DBuf<int32_t> tiles;  // tile row starts, ntiles + 1
  DBuf<int32_t> tnz;    // tile entry starts (ptr[tiles[t]]), ntiles + 1
  // workgroup tiles: k_spmv_stream sums every tile's rows one per lane (no
  // tile averages over 8 entries per row across more than 128 rows, none is
  // one long row), the condition of its delta-coded column format
  bool lane1 = false;
};

// SpMV plan of one CSR matrix: CSR-stream tiles (kind 0), L lanes per row
// (kind 1) or merge-path segments (kind 2).  Kinds 0 and 1 keep workgroup
// tiles (k_spmv_stream, the default: the faster kind for both value sizes on
// MI355X) and/or wavefront tiles (k_spmv_wave); the k-skip two-input SpMV
// uses the FP64 choice.  For kind 2, `seg` holds the segment start points
// (row, entry) of the merge path.
struct Plan {
  int kind = 0;
  bool wg64 = true, wg32 = true;  // tile set of the FP64 / FP32 SpMV
  int L = 8;
  TileSet wave, wg, seg;
  DBuf<double> carry;  // merge path: carries and heads, 2 x 2 values per segment (FP64 or FP32)
};

template <class T>
struct MatView {
  int64_t n = 0, nnz = 0;
  const int32_t* ptr = nullptr;
  const int32_t* col = nullptr;
  // 16-bit column indices, when the matrix has them and this view's SpMV
  // reads them (workgroup CSR-stream tiles): int16 offsets col - row (banded
  // matrices) or delta-coded columns (see k_spmv_stream); ib = bytes per
  // column index that the SpMV reads
  const int16_t* col16 = nullptr;
  const uint16_t* cold = nullptr;
  int ib = 4;
  const T* val = nullptr;
  int kind = 0;
  bool wg_tiles = false;
  int L = 8;
  int64_t ntiles = 0;
  const int32_t* tiles = nullptr;
  const int32_t* tnz = nullptr;
  T* carry = nullptr;
  // column panels (kind 3): np sub-matrices over all rows, each holding the
  // entries whose columns fall in one slice of x; pacc holds row partials
  int np = 0;
  const MatView* panel = nullptr;
  T* pacc = nullptr;
  void set_plan(const Plan& pl) {
    kind = pl.kind;
    L = pl.L;
    wg_tiles = sizeof(T) == 8 ? pl.wg64 : pl.wg32;
    const TileSet& ts = kind == 2 ? pl.seg : wg_tiles ? pl.wg : pl.wave;
    ntiles = ts.ntiles;
    tiles = ts.tiles.p;
    tnz = ts.tnz.p;
    carry = reinterpret_cast<T*>(pl.carry.p);
  }
};

// Column panels of one CSR matrix (A or A^T).  On matrices whose rows gather
// x from all over (random columns), x does not fit an XCD's L2 and every
// gather is a trip to the MALL.  Split into np panels by column, each SpMV
// pass gathers from a slice of x that the L2s hold; the passes add their row
// sums into pacc and the last one applies the epilogue.  Panel p's entries
// are a CSR over all n rows (ptr from 0), stored from off[p] (a multiple of
// four entries, so chunk loads stay 16-byte aligned); perm maps each panel
// entry to its entry of A's value array.
struct PanelSet {
  int np = 0;
  DBuf<int32_t> ptr, col, perm;  // np * (n + 1); panel entries; their A entries
  DBuf<double> val;
  DBuf<float> val32;
  std::vector<int64_t> off;  // np + 1
  std::vector<Plan> plans;
  std::vector<MatView<double>> v64;
  std::vector<MatView<float>> v32;
  DBuf<double> acc64;
  DBuf<float> acc32;
};

}  // namespace ks

struct ks_mat {
  ks_ctx* ctx = nullptr;
  int64_t n = 0, nnz = 0;
  uint32_t flags = 0;
  ks::DBuf<int32_t> ptr, col;
  ks::DBuf<int16_t> col16, tcol16;  // col - row of A and A^T (banded matrices)
  ks::DBuf<uint16_t> cold, tcold;   // delta-coded columns of A and A^T
  ks::DBuf<double> val;
  ks::DBuf<float> val32;
  bool has_t = false;
  ks::DBuf<int32_t> tptr, tcol, tperm;
  ks::DBuf<double> tval;
  ks::DBuf<float> tval32;
  ks::Plan plan, tplan;
  ks::PanelSet panels, tpanels;
  int64_t max_row = 0, max_trow = 0;

  template <class T>
  ks::MatView<T> view(bool trans) const;
  // hands the int16 offsets to a view whose SpMV reads them
  template <class T>
  void set_col16(ks::MatView<T>& v, bool trans) const {
    if (v.kind != 0 || !v.wg_tiles) return;
    const int16_t* p = trans ? tcol16.p : col16.p;
    const uint16_t* q = trans ? tcold.p : cold.p;
    if (p) { v.col16 = p; v.ib = 2; }
    else if (q) { v.cold = q; v.ib = 2; }
  }
};

template <>
inline ks::MatView<double> ks_mat::view<double>(bool trans) const {
  ks::MatView<double> v;
  v.n = n; v.nnz = nnz;
  const ks::Plan& pl = trans ? tplan : plan;
  if (trans) { v.ptr = tptr.p; v.col = tcol.p; v.val = tval.p; }
  else { v.ptr = ptr.p; v.col = col.p; v.val = val.p; }
  v.set_plan(pl);
  const ks::PanelSet& ps = trans ? tpanels : panels;
  if (ps.np) { v.np = ps.np; v.panel = ps.v64.data(); v.pacc = ps.acc64.p; v.kind = 3; }
  set_col16(v, trans);
  return v;
}
template <>
inline ks::MatView<float> ks_mat::view<float>(bool trans) const {
  ks::MatView<float> v;
  v.n = n; v.nnz = nnz;
  const ks::Plan& pl = trans ? tplan : plan;
  if (trans) { v.ptr = tptr.p; v.col = tcol.p; v.val = tval32.p; }
  else { v.ptr = ptr.p; v.col = col.p; v.val = val32.p; }
  v.set_plan(pl);
  const ks::PanelSet& ps = trans ? tpanels : panels;
  if (ps.np && ps.val32.p) { v.np = ps.np; v.panel = ps.v32.data(); v.pacc = ps.acc32.p; v.kind = 3; }
  set_col16(v, trans);
  return v;
}

namespace ks {

inline RedArgs red(ks_ctx* c) { return RedArgs{c->part.p}; }
inline StatArgs stat(ks_ctx* c, int cls, double bytes) {
  return StatArgs{c->stats.p, c->stats.p + KS_K_NCLASS, cls,
                  (unsigned long long)(bytes > 0 ? bytes : 0)};
}

inline int grid_map(int64_t n) {
  int64_t g = (n + (int64_t)kBlock * 4 - 1) / ((int64_t)kBlock * 4);
  if (g < 1) g = 1;
  if (g > kMaxGrid) g = kMaxGrid;
  if (g >= 64) g &= ~7;
  return (int)g;
}
inline int grid_rows(int64_t nrb) {
  int64_t g = nrb < kMaxGrid ? nrb : kMaxGrid;
  if (g < 1) g = 1;
  if (g >= 64) g &= ~7;
  return (int)g;
}

// Workgroups of `kernel` that are resident on the whole device at once.  The
// grid-stride kernels are launched with at most this many workgroups: a grid
// larger than one resident wave leaves a second, partly filled wave of
// workgroups (a tail) that runs after the rest.
template <class K>
int resident(ks_ctx* c, K kernel, size_t dyn_lds = 0) {
  const auto key = std::make_pair(reinterpret_cast<const void*>(kernel), dyn_lds);
  auto it = c->occ.find(key);
  if (it != c->occ.end()) return it->second;
  int per_cu = 0;
  KS_HIPCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel), kBlock,
                                                         dyn_lds));
  const int r = std::max(1, per_cu) * std::max(1, c->num_cus);
  c->occ.emplace(key, r);
  return r;
}
// grid_map / grid_rows, capped at one resident wave of `kernel`
template <class K>
int grid_map(ks_ctx* c, K kernel, int64_t n, size_t dyn_lds = 0) {
  return grid_rows(std::min<int64_t>(grid_map(n), resident(c, kernel, dyn_lds)));
}
template <class K>
int grid_rows(ks_ctx* c, K kernel, int64_t nrb, size_t dyn_lds = 0) {
  return grid_rows(std::min<int64_t>(nrb, resident(c, kernel, dyn_lds)));
}

// Launches kernel<<<G, kBlock, shmem>>> on the context's stream.  When
// kernel class `cls` is profiled, the launch carries a start/stop event pair
// that the runtime sets to the kernel's own execution interval (the
// dispatch's begin and end timestamps, as in a profiler's kernel trace), so
// the per-class times exclude the gaps between kernels.
template <class... P, class... A>
void klaunch(ks_ctx* c, int cls, void (*kernel)(P...), int G, size_t shmem, A&&... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->prof.pair(cls, e0, e1))
    hipExtLaunchKernelGGL(kernel, dim3(G), dim3(kBlock), (uint32_t)shmem, c->stream, e0, e1, 0u, P(args)...);
  else
    kernel<<<G, kBlock, shmem, c->stream>>>(P(args)...);
  KS_HIPCK(hipGetLastError());
}

// Finisher of a reduction whose producer ran on G workgroups.
template <int NV, class F>
void launch_finish(ks_ctx* c, int G, const F& f, Gate g, const double* part = nullptr) {
  k_finish<NV, F><<<1, kBlock, 0, c->stream>>>(G, part ? part : c->part.p, f, g);
  KS_HIPCK(hipGetLastError());
}

template <int NV, class Op>
void launch_map(ks_ctx* c, int64_t n, const Op& op, Gate g, int cls, double bytes) {
  const int G = grid_map(c, k_map<NV, Op>, n);
  klaunch(c, cls, k_map<NV, Op>, G, 0, n, op, g, red(c), stat(c, cls, bytes));
  if constexpr (NV > 0) launch_finish<NV>(c, G, op, g);
}
// k_map with its partials written to ra and no finisher: a later kernel
// forms the totals (partial_totals).  Returns the number of partials.
template <int NV, class Op>
int launch_map_partials(ks_ctx* c, int64_t n, const Op& op, Gate g, int cls, double bytes, RedArgs ra) {
  const int G = grid_map(c, k_map<NV, Op>, n);
  klaunch(c, cls, k_map<NV, Op>, G, 0, n, op, g, ra, stat(c, cls, bytes));
  return G;
}

template <class Op>
void launch_scalar(ks_ctx* c, const Op& op, Gate g) {
  k_scalar<Op><<<1, 1, 0, c->stream>>>(op, g, stat(c, KS_K_SCALAR, 0));
  KS_HIPCK(hipGetLastError());
}

// SpMV bytes: values + column indices (ib bytes each) + row pointers + x read
// once + y written.
template <class T>
inline double spmv_bytes(int64_t n, int64_t nnz, int ib = 4) {
  return (double)(sizeof(T) + ib) * nnz + 4.0 * (n + 1) + 2.0 * sizeof(T) * n;
}

template <class T, int L, int NV, class Epi>
int spmv_L(ks_ctx* c, const MatView<T>& A, const T* x, const Epi& epi, Gate g, int cls, double bytes) {
  constexpr int RPB = kBlock / L;
  const int64_t nrb = (A.n + RPB - 1) / RPB;
  const int G = grid_rows(c, k_spmv<T, L, NV, Epi>, nrb);
  klaunch(c, cls, k_spmv<T, L, NV, Epi>, G, 0, A.n, A.ptr, A.col, A.val, x, epi, g, red(c), stat(c, cls, bytes));
  return G;
}

// The workgroup CSR-stream SpMV in the view's column index format, on G
// workgroups.  G is always the int32 kernel's grid (grid_rows of that
// instantiation): the grid, and so the reduction partials and their
// summation order, is the same in every index format, so a solve's results
// do not depend on the format (the 16-bit kernels, which need fewer
// registers, may have room for more workgroups than they are given).
template <class T, int NX, int NV, class Epi>
void launch_stream(ks_ctx* c, int cls, const MatView<T>& A, int G, const T* x0, const T* x1, const Epi& epi, Gate g,
                   double bytes) {
  if (A.col16)
    klaunch(c, cls, k_spmv_stream<T, NX, NV, Epi, int16_t>, G, 0, A.n, A.ntiles, A.tiles, A.tnz, A.ptr, A.col16,
            A.val, x0, x1, epi, g, red(c), stat(c, cls, bytes));
  else if (A.cold)
    klaunch(c, cls, k_spmv_stream<T, NX, NV, Epi, uint16_t>, G, 0, A.n, A.ntiles, A.tiles, A.tnz, A.ptr, A.cold,
            A.val, x0, x1, epi, g, red(c), stat(c, cls, bytes));
  else
    klaunch(c, cls, k_spmv_stream<T, NX, NV, Epi>, G, 0, A.n, A.ntiles, A.tiles, A.tnz, A.ptr, A.col, A.val, x0, x1,
            epi, g, red(c), stat(c, cls, bytes));
}

template <class T, int NX, int NV, class Epi>
int spmv_stream(ks_ctx* c, const MatView<T>& A, const T* x0, const T* x1, const Epi& epi, Gate g, int cls,
                double bytes) {
  int G;
  if (A.wg_tiles) {
    G = grid_rows(c, k_spmv_stream<T, NX, NV, Epi>, A.ntiles);
    launch_stream<T, NX, NV, Epi>(c, cls, A, G, x0, x1, epi, g, bytes);
  } else {
    G = grid_rows(c, k_spmv_wave<T, NX, NV, Epi>, (A.ntiles + kBlock / 64 - 1) / (kBlock / 64));
    klaunch(c, cls, k_spmv_wave<T, NX, NV, Epi>, G, 0, A.n, A.ntiles, A.tiles, A.tnz, A.ptr, A.col, A.val, x0, x1,
            epi, g, red(c), stat(c, cls, bytes));
  }
  return G;
}

// Merge-path SpMV: one pass over the segments, then the rows that span
// segments.  Returns the number of reduction partials written (the main
// kernel's workgroups, then the fix-up kernel's).
template <class T, int NX, int NV, class Epi>
int spmv_merge(ks_ctx* c, const MatView<T>& A, const T* x0, const T* x1, const Epi& epi, Gate g, int cls,
               double bytes) {
  const int64_t nslot = (A.ntiles + kBlock / 64 - 1) / (kBlock / 64);
  c->prof.begin(c->stream, cls);
  int G = grid_rows(c, k_spmv_merge<T, NX, NV, Epi>, nslot);
  if (G > kMaxGrid - kMergeFixGrid) G = (kMaxGrid - kMergeFixGrid) & ~7;
  k_spmv_merge<T, NX, NV, Epi><<<G, kBlock, 0, c->stream>>>(A.n, A.ntiles, A.tiles, A.tnz, A.ptr, A.col, A.val, x0,
                                                            x1, A.carry, epi, g, red(c), stat(c, cls, bytes));
  KS_HIPCK(hipGetLastError());
  int64_t gf = (A.ntiles + kBlock - 1) / kBlock;
  const int Gf = (int)(gf < 1 ? 1 : gf > kMergeFixGrid ? kMergeFixGrid : gf);
  k_merge_fix<T, NX, NV, Epi><<<Gf, kBlock, 0, c->stream>>>(A.n, A.ntiles, A.tiles, A.tnz, A.ptr, A.carry, epi, g,
                                                            RedArgs{red(c).part + G});
  c->prof.end(c->stream);
  KS_HIPCK(hipGetLastError());
  return G + Gf;
}

// One SpMV launch on A's own layout (no finisher); returns the grid size.
template <class T, int NV, class Epi>
int spmv_dispatch(ks_ctx* c, const MatView<T>& A, const T* x, const Epi& epi, Gate g, int cls, double bytes) {
  if (A.kind == 0) return spmv_stream<T, 1, NV>(c, A, x, x, epi, g, cls, bytes);
  if (A.kind == 2) return spmv_merge<T, 1, NV>(c, A, x, x, epi, g, cls, bytes);
  switch (A.L) {
    case 2: return spmv_L<T, 2, NV>(c, A, x, epi, g, cls, bytes);
    case 4: return spmv_L<T, 4, NV>(c, A, x, epi, g, cls, bytes);
    case 8: return spmv_L<T, 8, NV>(c, A, x, epi, g, cls, bytes);
    case 16: return spmv_L<T, 16, NV>(c, A, x, epi, g, cls, bytes);
    case 32: return spmv_L<T, 32, NV>(c, A, x, epi, g, cls, bytes);
    default: return spmv_L<T, 64, NV>(c, A, x, epi, g, cls, bytes);
  }
}

// Column-panel passes: every pass but the last stores (first) or adds its
// row sums into pacc; the last adds pacc to its own and runs the epilogue.
template <class T>
struct EpiPanelAcc {
  T* acc;
  int first;
  __device__ void prep() {}
  __device__ void row(int64_t i, T s, T, double*) { acc[i] = first ? s : acc[i] + s; }
  __device__ void fin(const double*) const {}
};
template <class T, class Epi>
struct EpiPanelLast {
  Epi e;
  const T* acc;
  __device__ void prep() { e.prep(); }
  __device__ void row(int64_t i, T s, T s1, double* a) { e.row(i, acc[i] + s, s1, a); }
  __device__ void fin(const double* s) const { e.fin(s); }
};

// y-side epilogue of SpMV plus NV reduction values (finished by k_finish).
// With finish false no finisher is launched (a later kernel forms the
// totals of the G partials, G returned).
template <class T, int NV, class Epi>
int launch_spmv(ks_ctx* c, const MatView<T>& A, const T* x, const Epi& epi, Gate g, double extra_bytes = 0,
                bool finish = true) {
  const int cls = sizeof(T) == 8 ? KS_K_SPMV64 : KS_K_SPMV32;
  const double bytes = spmv_bytes<T>(A.n, A.nnz, A.ib) + extra_bytes;
  int G;
  if (A.np > 0) {
    // the bytes of the whole SpMV are counted with the last pass
    for (int p = 0; p + 1 < A.np; ++p)
      spmv_dispatch<T, 0>(c, A.panel[p], x, EpiPanelAcc<T>{A.pacc, p == 0}, g, cls, 0.0);
    G = spmv_dispatch<T, NV>(c, A.panel[A.np - 1], x, EpiPanelLast<T, Epi>{epi, A.pacc}, g, cls, bytes);
  } else {
    G = spmv_dispatch<T, NV>(c, A, x, epi, g, cls, bytes);
  }
  if constexpr (NV > 0)
    if (finish) launch_finish<NV>(c, G, epi, g);
  return G;
}

template <class T>
struct EpiStore;
template <class T>
struct EpiStore2;

// y1 = A x1 and y2 = A x2 in one pass over the matrix (always CSR-stream,
// with the plan's tiles).
template <class T>
void launch_spmv2(ks_ctx* c, const MatView<T>& A, const T* x1, const T* x2, T* y1, T* y2, Gate g) {
  if (A.np > 0) {  // column panels: two single-vector SpMVs
    launch_spmv<T, 0>(c, A, x1, EpiStore<T>{y1}, g);
    launch_spmv<T, 0>(c, A, x2, EpiStore<T>{y2}, g);
    return;
  }
  const double bytes = spmv_bytes<T>(A.n, A.nnz, A.ib) + 2.0 * sizeof(T) * A.n;
  if (A.kind == 2) spmv_merge<T, 2, 0>(c, A, x1, x2, EpiStore2<T>{y1, y2}, g, KS_K_BASIS, bytes);
  else spmv_stream<T, 2, 0>(c, A, x1, x2, EpiStore2<T>{y1, y2}, g, KS_K_BASIS, bytes);
}

// CSR-stream tiles from host row pointers: greedy runs of at most max_rows
// rows and max_nnz entries; a longer row is a tile of its own.
inline std::vector<int32_t> build_tiles(const int32_t* hp, int64_t n, int64_t max_rows, int64_t max_nnz) {
  std::vector<int32_t> t;
  t.reserve(n / 32 + 2);
  t.push_back(0);
  int64_t r = 0;
  while (r < n) {
    const int64_t start = r, j0 = hp[r];
    if (hp[r + 1] - j0 > max_nnz) {
      ++r;
    } else {
      while (r < n && r - start < max_rows && hp[r + 1] - j0 <= max_nnz) ++r;
    }
    t.push_back((int32_t)r);
  }
  return t;
}

// Merge-path segment starts: the path point (row i, entry j) on each
// diagonal d = s * items (Merrill & Garland's MergePathSearch with
// A[i] = ptr[i + 1], B[j] = j: a row end is taken before entry j when
// ptr[i + 1] <= j).
inline void merge_segments(const int32_t* hp, int64_t n, int64_t nnz, int64_t items, std::vector<int32_t>& segr,
                           std::vector<int32_t>& segj) {
  const int64_t total = n + nnz, nseg = total > 0 ? (total + items - 1) / items : 1;
  segr.resize(nseg + 1);
  segj.resize(nseg + 1);
  for (int64_t s = 0; s <= nseg; ++s) {
    const int64_t d = std::min(s * items, total);
    int64_t lo = std::max<int64_t>(d - nnz, 0), hi = std::min<int64_t>(d, n);
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (hp[mid + 1] <= d - mid - 1) lo = mid + 1;
      else hi = mid;
    }
    segr[s] = (int32_t)lo;
    segj[s] = (int32_t)(d - lo);
  }
}

inline void upload_tiles(ks_ctx* c, TileSet& ts, const std::vector<int32_t>& t, const std::vector<int32_t>& z) {
  ts.ntiles = (int64_t)t.size() - 1;
  ts.tiles.alloc(t.size());
  ts.tnz.alloc(t.size());
  KS_HIPCK(hipMemcpyAsync(ts.tiles.p, t.data(), sizeof(int32_t) * t.size(), hipMemcpyHostToDevice, c->stream));
  KS_HIPCK(hipMemcpyAsync(ts.tnz.p, z.data(), sizeof(int32_t) * z.size(), hipMemcpyHostToDevice, c->stream));
  KS_HIPCK(hipStreamSynchronize(c->stream));
}

inline void tile_set(ks_ctx* c, TileSet& ts, const int32_t* hp, int64_t n, int64_t rows, int64_t nnz_max) {
  std::vector<int32_t> t = build_tiles(hp, n, rows, nnz_max), z(t.size());
  for (size_t i = 0; i < t.size(); ++i) z[i] = hp[t[i]];
  // k_spmv_stream's lanes per row for a tile of nr rows and cnt entries is 1
  // unless 2 nr <= kBlock and 8 nr < cnt; a tile over kTileNnz entries is
  // one long row
  ts.lane1 = true;
  for (size_t i = 0; i + 1 < t.size() && ts.lane1; ++i) {
    const int64_t nr = t[i + 1] - t[i], cnt = (int64_t)z[i + 1] - z[i];
    if (cnt > kTileNnz || (2 * nr <= kBlock && 8 * nr < cnt)) ts.lane1 = false;
  }
  upload_tiles(c, ts, t, z);
}

// tiles: 0 or 1 workgroup tiles (the default), 2 wavefront tiles.
inline void make_plan(ks_ctx* c, Plan& pl, const int32_t* hp, int64_t n, int64_t nnz, bool vector, int tiles,
                      bool merge = false) {
  pl.L = 2;
  const double mean = n > 0 ? (double)nnz / (double)n : 1.0;
  while (pl.L < 64 && pl.L < mean) pl.L *= 2;
  pl.kind = merge ? 2 : vector ? 1 : 0;
  if (merge) {
    std::vector<int32_t> t, z;
    merge_segments(hp, n, nnz, kMergeItems, t, z);
    pl.carry.alloc(4 * (t.size() - 1));
    upload_tiles(c, pl.seg, t, z);
    return;
  }
  pl.wg64 = pl.wg32 = tiles != 2;
  if (!pl.wg64 || !pl.wg32) tile_set(c, pl.wave, hp, n, kWaveRows, kWaveNnz);
  if (pl.wg64 || pl.wg32) tile_set(c, pl.wg, hp, n, kTileRows, kTileNnz);
}

// Lanes per row: smallest power of two >= mean row length, in [2, 64].
inline int choose_lanes(int64_t n, int64_t nnz) {
  const double mean = n > 0 ? (double)nnz / (double)n : 1.0;
  int L = 2;
  while (L < 64 && L < mean) L *= 2;
  return L;
}

// ---------------------------------------------------------- epilogues ---
template <class T>
struct EpiStore {
  T* y;
  __device__ void prep() {}
  __device__ void row(int64_t i, T s, T, double*) { y[i] = s; }
  __device__ void fin(const double*) const {}
};
// y = A x streamed out with non-temporal stores: the standalone SpMV, whose
// result goes back to the caller.  Inside the solvers the next kernel reads y
// at once and plain stores (EpiStore) keep it in the memory-side cache.
template <class T>
struct EpiStream {
  T* y;
  __device__ void prep() {}
  __device__ void row(int64_t i, T s, T, double*) { __builtin_nontemporal_store(s, y + i); }
  __device__ void fin(const double*) const {}
};
// y = (A x) * r (r: device FP64 reciprocal of the basis norm, formed by the
// norm finisher; rounded to T once per thread and held in scalar registers,
// as it is uniform)
template <class T>
struct EpiDiv {
  T* y;
  const double* r;
  T inv;
  __device__ void prep() {
    const T v = (T)*r;
    if constexpr (sizeof(T) == 4) {
      inv = __builtin_bit_cast(T, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
    } else {
      const uint64_t b = __builtin_bit_cast(uint64_t, v);
      const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
      const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
      inv = __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    }
  }
  __device__ T value(T s) const { return s * inv; }
  __device__ void row(int64_t i, T s, T, double*) { y[i] = value(s); }
  __device__ void fin(const double*) const {}
};
// sqrt of the sum of the ng <= kMaxGrid partials p[] (a norm whose update
// kernel deferred its finisher), formed by the whole workgroup in a fixed
// order: every workgroup, and every kernel, that calls it gets the same
// bits.  All threads must call it; all return the value.
__device__ inline double norm_from_partials(const double* __restrict__ p, int ng) {
  static_assert(kMaxGrid == 8 * kBlock && kBlock == 256, "eight partials per thread, four wavefronts");
  __shared__ double sw[kBlock / 64];
  double t[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) t[u] = p[threadIdx.x + u * kBlock];  // in bounds: kMaxGrid entries
#pragma unroll
  for (int u = 0; u < 8; ++u)
    if ((int)threadIdx.x + u * kBlock >= ng) t[u] = 0.0;
  double q = ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
  q = wave_sum(q);
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = q;
  __syncthreads();
  return sqrt((sw[0] + sw[1]) + (sw[2] + sw[3]));
}

// EpiDiv plus the projection h_v = (y, V_v), v < nv (nv <= kMaxNV), formed
// by the workgroup CSR-stream kernel over its own rows after the SpMV.
// With np set, the scale is taken from the previous GMRES step's deferred
// norm: 1 / sqrt(sum of the ng partials np[] its update kernel left), formed
// by every workgroup with norm_from_partials; that step's norm finisher
// itself runs in this launch's projection finisher (FinNormProj).  With
// neither np nor r set the rows are not scaled.
template <class T>
struct EpiDivProj : EpiDiv<T> {
  const T* V;
  int64_t ldv;
  int nv;
  const double* np;
  int ng;
  static constexpr bool kProj = true;
  __device__ void prep() {
    if (np) this->inv = (T)(1.0 / norm_from_partials(np, ng));
    else if (this->r) EpiDiv<T>::prep();
    else this->inv = T(1);  // no scale: y = A x exactly
  }
};
template <class T>
struct EpiStore2 {
  T *y0, *y1;
  __device__ void prep() {}
  __device__ void row(int64_t i, T s0, T s1, double*) { y0[i] = s0; y1[i] = s1; }
  __device__ void fin(const double*) const {}
};

// y = b - A x (b FP64), acc += y^2, Fin on the total.
template <class T, class Fin>
struct EpiResid {
  const double* b;
  T* y;
  Fin f;
  __device__ void prep() {}
  __device__ void row(int64_t i, T s, T, double* acc) {
    const T r = (T)(b[i] - (double)s);
    y[i] = r;
    acc[0] += (double)r * (double)r;
  }
  __device__ void fin(const double* s) const { f(s); }
};

// y = A x, acc += d . y, Fin on the total.
template <class T, class Fin>
struct EpiDot {
  const T* d;
  T* y;
  Fin f;
  __device__ void prep() {}
  __device__ void row(int64_t i, T s, T, double* acc) {
    y[i] = s;
    acc[0] += (double)d[i] * (double)s;
  }
  __device__ void fin(const double* s) const { f(s); }
};

struct FinNone {
  __device__ void operator()(const double*) const {}
};
// writes sum (or its square root) to a device scalar
struct FinStore {
  double* out;
  int sqrt_it;
  __device__ void operator()(const double* s) const { *out = sqrt_it ? sqrt(s[0]) : s[0]; }
};

}  // namespace ks
