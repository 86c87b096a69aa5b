// ks_kernels.h - HIP kernels for gfx950 (CDNA4, wave64).
//
// Every kernel here is HBM-bound: the CSR SpMV reads (s_v+4)*nnz bytes of
// matrix per call, the BLAS-1 / Gram-Schmidt kernels stream whole vectors.
// Design points:
//  * one generic grid-stride map/reduce kernel (k_map) carries every fused
//    elementwise update + dot product of the solvers; reductions are
//    deterministic two-stage sums (fixed grid, fixed per-thread element
//    sets, fixed shuffle tree; a one-workgroup finisher, k_finish, sums the
//    per-workgroup partials in a fixed order and runs the scalar step);
//  * CSR SpMV: CSR-stream tiles (k_spmv_stream per workgroup, k_spmv_wave
//    per wavefront) that stage values and column indices in LDS and sum one
//    row per lane; a merge-path variant (k_spmv_merge + k_merge_fix); and
//    a lanes-per-row kernel (k_spmv).  Work is assigned to XCDs in
//    contiguous ranges so x gathers hit the XCD's own L2.  On workgroup
//    tiles an epilogue can also take work that needs no global
//    synchronisation after the SpMV: the GMRES projection V^T w over the
//    workgroup's own rows (kProj);
//  * solver control flow stays on the device: kernels take a Gate and return
//    at once when the solve (or the enclosing outer step) has stopped, so
//    the host can enqueue many iterations between synchronisations;
//  * scalar recurrences (alpha, beta, Givens rotations, k-skip updates) run
//    in the finaliser of the reduction that produces their inputs, or in a
//    one-thread kernel, in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace ks {

constexpr int kBlock = 256;       // threads per workgroup (4 wavefronts)
constexpr int kMaxGrid = 2048;    // workgroups of a reduction grid (8/CU at 256 CUs)
constexpr int kMaxNV = 64;        // reduction values per kernel
constexpr int kXcd = 8;           // XCDs per MI355X; workgroup b runs on XCD b % 8

// ---------------------------------------------------------------- control --
// Device-resident state of one solve level (outer or inner).
struct Ctl {
  int stop;        // 1: no further Krylov steps (converged / limit / breakdown)
  int done;        // 1: final assembly done (GMRES family)
  int status;      // 0 converged, 2 not converged, 3 breakdown
  int ksz;         // GMRES: steps completed in the cycle that stopped
  long long count; // iteration counter (meaning per solver)
  long long nhist;
  long long hist_cap;
  long long maxloop;
  long long steps;
  long long tsteps;  // steps over all solves of this level (never reset)
  double eps, bnorm, err;
  double rr, alpha, beta, gamma, dotv, vscale, rz;
  double rrs[2];   // CG: rr by iteration parity (kernels with a folded finisher)
  double* hist;    // (idx, err) pairs, device
  int* host_stop;  // host-mapped flag mirror of `stop` (nullable)
  int R, ld;       // GMRES restart and H row stride
  double *H, *e, *cs, *sn, *y;
  double* vn;      // GMRES: norms of the unnormalised basis vectors (R)
  double* vr;      // GMRES: 1 / vn (R), the SpMV epilogue's scale
  int kskip, fix, compat;
  int xzero;       // GMRES: x is still the zero it started from (set at a cycle start)
  double *delta, *eta, *zeta, *rho, *phi;
};

__device__ inline void ctl_stop(Ctl* c, int status) {
  c->stop = 1;
  c->status = status;
  if (c->host_stop) *(volatile int*)c->host_stop = 1;
}
__device__ inline void ctl_hist(Ctl* c, double idx, double err) {
  c->err = err;
  if (c->hist && c->nhist < c->hist_cap) {
    c->hist[2 * c->nhist] = idx;
    c->hist[2 * c->nhist + 1] = err;
  }
  c->nhist++;
}
__device__ inline bool finite(double v) { return __builtin_isfinite(v); }

// A kernel is skipped when either flag is set.
struct Gate {
  const int* a;
  const int* b;
  __device__ bool off() const { return (a && *a) || (b && *b); }
};

// Reduction scratch of the context: partial sums, value-major with stride
// kMaxGrid: part[i * kMaxGrid + workgroup].
struct RedArgs {
  double* part;
};

// Executed-launch statistics per kernel class (counted by workgroup 0).
struct StatArgs {
  unsigned long long* launches;
  unsigned long long* bytes;
  int cls;
  unsigned long long nbytes;
  __device__ void count() const {
    if (launches && blockIdx.x == 0 && threadIdx.x == 0) {
      atomicAdd(launches + cls, 1ull);
      atomicAdd(bytes + cls, nbytes);
    }
  }
};

// ----------------------------------------------------------- reductions --
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Stores this workgroup's sums of acc[0..NV) as its partials.  The totals
// are formed by k_finish, launched right after the producing kernel: the
// kernel boundary orders the partials before their use, so no device-scope
// fence (which on gfx950 writes back and invalidates the XCD's L2) is needed.
template <int NV>
__device__ void block_store(double (&acc)[NV], RedArgs ra) {
  __shared__ double sw[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int nw = kBlock / 64;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double v = wave_sum(acc[i]);
    if (lane == 0) sw[i][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < nw; ++k) s += sw[threadIdx.x][k];
    ra.part[(size_t)threadIdx.x * kMaxGrid + blockIdx.x] = s;
  }
}

// Totals of the G partials of each of the NV values, then f.fin(sums) on one
// thread.  TPV = kBlock / NVP threads share one value (NVP = NV rounded up to
// a power of two, at least 4): each sums a fixed strided subset in rounds of
// sixteen (then four, then one) independent loads (the partials come from
// other XCDs' writes, so each round is a full memory round trip: few rounds
// keep the finisher short), then a fixed tree over the sixteen sums and a
// shuffle tree inside the TPV-lane group - a fixed order, so the totals are
// deterministic.
template <class F, class = void>
struct HasPre : std::false_type {};
template <class F>
struct HasPre<F, std::void_t<decltype(std::declval<F&>().pre())>> : std::true_type {};

// The totals themselves, by the whole workgroup (every thread must call it;
// all get the same LDS array back): k_finish's step, also taken by kernels
// that fold a finisher into their own start, with the same bits.
template <int NV>
__device__ __forceinline__ const double* partial_totals(int G, const double* __restrict__ part) {
  constexpr int NVP = NV <= 4 ? 4 : NV <= 8 ? 8 : NV <= 16 ? 16 : NV <= 32 ? 32 : 64;
  constexpr int TPV = kBlock / NVP;  // 4..64, a divisor of the wavefront
  constexpr int U = 16;
  __shared__ double sums[NVP];
  const int v = threadIdx.x / TPV, j = threadIdx.x % TPV;
  double a[U];
#pragma unroll
  for (int u = 0; u < U; ++u) a[u] = 0.0;
  if (v < NV) {
    const double* p = part + (size_t)v * kMaxGrid;
    int b = j;
    // full rounds of U loads, then of 4, then single loads; every load in a
    // round is unconditional, so the round's loads issue back to back
    for (; b + (U - 1) * TPV < G; b += U * TPV) {
      double t[U];
#pragma unroll
      for (int u = 0; u < U; ++u) t[u] = p[b + u * TPV];
#pragma unroll
      for (int u = 0; u < U; ++u) a[u] += t[u];
    }
    for (; b + 3 * TPV < G; b += 4 * TPV) {
      double t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) t[u] = p[b + u * TPV];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] += t[u];
    }
    for (; b < G; b += TPV) a[0] += p[b];
  }
#pragma unroll
  for (int w = U / 2; w > 0; w >>= 1)
#pragma unroll
    for (int u = 0; u < w; ++u) a[u] += a[u + w];
  double s = a[0];
#pragma unroll
  for (int off = TPV / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (j == 0) sums[v] = s;
  __syncthreads();
  return sums;
}

// F may have pre(), run by every thread before the totals (a workgroup-wide
// step of its own, such as norm_from_partials).
template <int NV, class F>
__global__ __launch_bounds__(kBlock) void k_finish(int G, const double* __restrict__ part, F f, Gate g) {
  if (g.off()) return;
  if constexpr (HasPre<F>::value) f.pre();
  const double* sums = partial_totals<NV>(G, part);
  if (threadIdx.x == 0) f.fin(sums);
}

// ------------------------------------------------------- generic map ----
// op.prep() once per thread (load scalars), op.elem(i, acc) per element,
// op.fin(sums) by the finisher (k_finish, launched after it) when NV > 0.
template <int NV, class Op>
__global__ __launch_bounds__(kBlock) void k_map(int64_t n, Op op, Gate g, RedArgs ra, StatArgs st) {
  if (g.off()) return;
  st.count();
  op.prep();
  double acc[NV > 0 ? NV : 1];
#pragma unroll
  for (int i = 0; i < (NV > 0 ? NV : 1); ++i) acc[i] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    op.elem(i, acc);
    op.elem(i + stride, acc);
    op.elem(i + 2 * stride, acc);
    op.elem(i + 3 * stride, acc);
  }
  for (; i < n; i += stride) op.elem(i, acc);
  if constexpr (NV > 0) block_store<NV>(acc, ra);
}

// One-thread kernel for scalar recurrences.
template <class Op>
__global__ void k_scalar(Op op, Gate g, StatArgs st) {
  if (g.off()) return;
  st.count();
  op.run();
}

// -------------------------------------------------------------- SpMV ----
// Row-block r of `nrb` for workgroup slot `it`: XCD-contiguous ranges.
// With nsub > 1 each workgroup is split into nsub slots (its wavefronts),
// slot `sub` of workgroup b taking every (slots per XCD)-th block of the
// XCD's range.
struct RowBlocks {
  int64_t nrb;
  __device__ void range(int64_t& first, int64_t& end, int64_t& step, int sub = 0, int nsub = 1) const {
    const int G = gridDim.x;
    if (G % kXcd == 0 && G >= 8 * kXcd) {
      const int xcd = blockIdx.x % kXcd, bi = blockIdx.x / kXcd, gx = G / kXcd;
      const int64_t per = (nrb + kXcd - 1) / kXcd;
      const int64_t lo = (int64_t)xcd * per;
      int64_t hi = lo + per;
      if (hi > nrb) hi = nrb;
      first = lo + (int64_t)bi * nsub + sub;
      end = hi;
      step = (int64_t)gx * nsub;
    } else {
      first = (int64_t)blockIdx.x * nsub + sub;
      end = nrb;
      step = (int64_t)G * nsub;
    }
  }
};

// y = A x with L lanes per row (L | 64), RPB = kBlock/L rows per row block.
// Epi: NV reduction values, prep(), row(row, sum, acc), fin(sums).
template <class T, int L, int NV, class Epi>
__global__ __launch_bounds__(kBlock) void k_spmv(int64_t n, const int32_t* __restrict__ ptr,
                                                 const int32_t* __restrict__ col,
                                                 const T* __restrict__ val, const T* __restrict__ x,
                                                 Epi epi, Gate g, RedArgs ra, StatArgs st) {
  if (g.off()) return;
  st.count();
  epi.prep();
  constexpr int RPB = kBlock / L;
  __shared__ int32_t sptr[RPB + 1];
  double acc[NV > 0 ? NV : 1];
#pragma unroll
  for (int i = 0; i < (NV > 0 ? NV : 1); ++i) acc[i] = 0.0;
  const int grp = threadIdx.x / L, lane = threadIdx.x % L;
  RowBlocks rbs{(n + RPB - 1) / RPB};
  int64_t rb, end, step;
  rbs.range(rb, end, step);
  for (; rb < end; rb += step) {
    const int64_t row0 = rb * RPB;
    __syncthreads();
    if (threadIdx.x <= RPB) {
      int64_t r = row0 + threadIdx.x;
      sptr[threadIdx.x] = ptr[r < n ? r : n];
    }
    __syncthreads();
    const int64_t row = row0 + grp;
    const int s = sptr[grp], e = sptr[grp + 1];
    T sum = 0;
    for (int j = s + lane; j < e; j += L) sum += val[j] * x[col[j]];
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, L);
    if (lane == 0 && row < n) epi.row(row, sum, (T)0, acc);
  }
  if constexpr (NV > 0) block_store<NV>(acc, ra);
}

// CSR-stream SpMV.  The rows are cut (on the host, once per matrix) into
// tiles of at most kTileRows rows and kTileNnz entries; a row longer than
// kTileNnz is a tile of its own.  tiles[t] / tnz[t] are the first row and
// first entry of tile t.  Per tile the workgroup (1) stages the tile's
// values and column indices in LDS (read from HBM coalesced, consecutive
// entries on consecutive lanes), then (2) sums each row from LDS with `tpr`
// threads per row (tpr = 1 for rows of up to 8 entries: a sequential sum in
// CSR order, the reference's blas::MtxVec_mult order, blas.cpp:56-62),
// gathering x as it goes.  With one row per lane, lane l gathers x for the
// j-th entry of row r0 + l, so on stencil-like matrices one gather
// instruction reads a contiguous run of x instead of the ~7 separate bands
// that consecutive entries touch.  The loop is software-pipelined:
// while tile t is summed, the bounds, row pointers, values and column
// indices of the workgroup's next tile are already in flight (eight
// independent loads per thread).  NX = 2 forms A x0 and A x1 from one pass
// over the matrix (the k-skip matrix-powers basis).
// Values and column indices are read in aligned chunks (16 bytes of values;
// int32 columns in chunks of as many entries, int16 offsets in 16-byte
// chunks of 8) and staged in LDS at their offsets from the aligned tile
// start a0 (j0 & ~3, or j0 & ~7 with int16 offsets) with aligned LDS stores.
// The chunks cover [a0, a0 + window), so a workgroup tile holds at most
// window - 7 entries; a chunk that would pass nnz is read entry by entry, so
// no load passes nnz.
constexpr int kTileWin = 2048;
constexpr int kTileNnz = kTileWin - 7;
constexpr int kTileRows = 256;
static_assert(kTileRows <= kBlock, "a workgroup tile row per thread");

template <class T>
__device__ inline T ld_stream(const T* p) { return __builtin_nontemporal_load(p); }

// a + v * x with the product rounded to T before the add (no FMA): the
// reference's `out += val * in` order (blas.cpp:56-62).
template <class T>
__device__ inline T mul_add_rn(T a, T v, T x) {
#pragma clang fp contract(off)
  return a + v * x;
}

template <class T>
struct Vec2 {
  typedef T type __attribute__((ext_vector_type(2)));
};
template <class T>
struct Vec4 {
  typedef T type __attribute__((ext_vector_type(4)));
};

// A chunk of W consecutive entries, loaded and stored as one vector.
template <class U, int W>
struct Chunk {
  typedef U type __attribute__((ext_vector_type(W)));
};

// One array's share of a tile for one thread: NTH threads cover a window of
// WIN entries from a0, thread t taking the W-entry chunks at a0 + W t +
// W NTH i, so each load instruction reads one chunk per lane from
// consecutive lanes.
template <class U, int NTH, int WIN, int W>
struct TileArr {
  static constexpr int kN = WIN / (W * NTH);
  U x[kN][W];
  // kFull: the whole window [a0, a0 + WIN) lies below nnz (uniform across
  // the tile), so no lane needs the entry-by-entry path; keeping that path
  // out of the common case keeps the chunk loads back to back.
  template <bool kFull>
  __device__ __forceinline__ void load(const U* __restrict__ p, int32_t a0, int32_t j1, int32_t nnz, int t) {
#pragma unroll
    for (int i = 0; i < kN; ++i) {
      const int32_t e = a0 + W * t + W * NTH * i;
      if (e < j1) {
        if (kFull || e + W - 1 < nnz) {
          const typename Chunk<U, W>::type q = ld_stream(reinterpret_cast<const typename Chunk<U, W>::type*>(p + e));
#pragma unroll
          for (int h = 0; h < W; ++h) x[i][h] = q[h];
        } else {
          // every slot written on both paths, so the array stays in registers
#pragma unroll
          for (int h = 0; h < W; ++h) x[i][h] = e + h < nnz ? ld_stream(p + e + h) : U(0);
        }
      }
    }
  }
  // chunks at their offsets from a0 (16-byte aligned LDS stores); slots
  // outside [j0, j1) are written too and never read
  __device__ __forceinline__ void stage(U* s, int32_t span, int t) const {
#pragma unroll
    for (int i = 0; i < kN; ++i) {
      const int32_t k = W * t + W * NTH * i;
      if (k < span) {
        typename Chunk<U, W>::type q;
#pragma unroll
        for (int h = 0; h < W; ++h) q[h] = x[i][h];
        *reinterpret_cast<typename Chunk<U, W>::type*>(s + k) = q;
      }
    }
  }
};

// One thread's share of a tile: values and column indices of the window
// [a0, a0 + WIN), a0 = a0of(j0), plus the row pointer of row r0 + t.  Chunks
// are 16 bytes of values (4 FP32 or 2 FP64 entries) and column indices:
// int32 columns, as many as the value chunk (16 or 8 bytes), or int16
// offsets col - row, 8 to a 16-byte chunk (see k_spmv_stream).
template <class T, int NTH, int WIN, class IX = int32_t>
struct QuadTile {
  static constexpr int W = 16 / (int)sizeof(T);
  static constexpr int WC = sizeof(IX) == 2 ? 8 : W;  // int16 offsets: 16-byte chunks of 8
  static constexpr int kA = WC > 4 ? 8 : 4;           // window start alignment (entries)
  __device__ static __forceinline__ int32_t a0of(int32_t j0) { return j0 & ~(kA - 1); }
  TileArr<T, NTH, WIN, W> v;
  TileArr<IX, NTH, WIN, WC> c;
  int32_t pr;
  __device__ __forceinline__ void load(const int32_t* __restrict__ ptr, const IX* __restrict__ col,
                                       const T* __restrict__ val, int32_t r0, int32_t r1, int32_t j0, int32_t j1,
                                       int32_t nnz, int t) {
    if (j1 - j0 > WIN - kA + 1) return;  // long row: loaded in place
    const int32_t a0 = a0of(j0);
    if (a0 + WIN <= nnz) {
      v.template load<true>(val, a0, j1, nnz, t);
      c.template load<true>(col, a0, j1, nnz, t);
    } else {
      v.template load<false>(val, a0, j1, nnz, t);
      c.template load<false>(col, a0, j1, nnz, t);
    }
    pr = t < r1 - r0 ? ptr[r0 + t] : 0;
  }
  __device__ __forceinline__ void stage(T* sv, IX* sc, int32_t j0, int32_t j1, int t) const {
    const int32_t span = j1 - a0of(j0);
    v.stage(sv, span, t);
    c.stage(sc, span, t);
  }
};
template <class T, class IX = int32_t>
using TileData = QuadTile<T, kBlock, kTileWin, IX>;

// Register budget of the CSR-stream kernels as waves/SIMD: FP32 64 VGPRs
// (8 waves) for a plain SpMV and 72 (7 waves) with a fused reduction, FP64
// 96 (5 waves) and 128 (4 waves) for the two-input SpMV - the largest
// occupancy each reaches without spilling.
template <class T, int NX, int NV>
constexpr int kStreamWaves = sizeof(T) == 4 ? (NV == 0 ? 8 : 7) : (NX == 2 ? 4 : 5);
// The workgroup kernel's FP32 budgets are one step larger (72 / 80 VGPRs):
// its scaled and projection epilogues take 66-69 VGPRs, and at 64 they spill.
template <class T, int NX, int NV>
constexpr int kWgWaves = sizeof(T) == 4 ? (NV == 0 ? 7 : 6) : kStreamWaves<T, NX, NV>;

// An epilogue with `kProj` also asks the workgroup kernel for the partial
// projections (y, V_v), v < nv, over the rows it produced (see
// k_spmv_stream).
template <class E, class = void>
struct HasProj : std::false_type {};
template <class E>
struct HasProj<E, std::void_t<decltype(E::kProj)>> : std::true_type {};

// IX = int16_t: the column index arrives as the offset col - row (every
// offset of the matrix in [-32767, 32767]; stencil and other banded
// matrices), half the index bytes of int32 columns.
// IX = uint16_t: delta-coded columns (matrices whose offsets do not fit
// int16 but whose rows are sorted with gaps below 65536, such as 3-D stencils
// with planes over 32767 points): a row's first entry holds col - row +
// kDeltaBias and every later one the gap to the previous column, decoded by
// the one lane that sums the row (the host builds this format only when
// every tile sums its rows one per lane and no row is a tile of its own).
// In every format the row sums are the same: entry for entry, the same x is
// gathered in the same order.
constexpr int32_t kDeltaBias = 49152;
template <class T, int NX, int NV, class Epi, class IX = int32_t>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWgWaves<T, NX, NV>))) void k_spmv_stream(int64_t n, int64_t ntiles, const int32_t* __restrict__ tiles,
                                                        const int32_t* __restrict__ tnz,
                                                        const int32_t* __restrict__ ptr,
                                                        const IX* __restrict__ col,
                                                        const T* __restrict__ val, const T* __restrict__ x0,
                                                        const T* __restrict__ x1, Epi epi, Gate g, RedArgs ra,
                                                        StatArgs st) {
  if (g.off()) return;
  st.count();
  constexpr bool kRel = std::is_same<IX, int16_t>::value;
  constexpr bool kDelta = std::is_same<IX, uint16_t>::value;
  // With the projection epilogue, the y of the workgroup's first kYT tiles
  // is kept in LDS (tile s at sy[s][row - (r0 & ~3)], 16-byte aligned row
  // chunks): the projection reads it there instead of from L2/HBM, and it
  // goes to HBM in one burst per tile during the projection rather than
  // between the matrix reads.  kYT is what fits beside the tile buffers at
  // the kernel's occupancy (7 FP32 / 5 FP64 workgroups per CU in 160 KB).
  constexpr int kYT = !HasProj<Epi>::value ? 0 : sizeof(T) == 4 ? (sizeof(IX) == 2 ? 9 : 4) : (sizeof(IX) == 2 ? 4 : 2);
  constexpr int kYW = kTileRows + 4;
  __shared__ __align__(16) T sv[kTileWin];
  __shared__ __align__(16) IX sc[kTileWin];
  __shared__ __align__(16) T sy[kYT > 0 ? kYT : 1][kYT > 0 ? kYW : 1];
  __shared__ int32_t sptr[kTileRows + 1];
  __shared__ double slong[NX][kBlock / 64];
  double acc[NV > 0 ? NV : 1];
#pragma unroll
  for (int i = 0; i < (NV > 0 ? NV : 1); ++i) acc[i] = 0.0;
  const int tid = threadIdx.x;
  RowBlocks rbs{ntiles};
  int64_t t, end, step;
  rbs.range(t, end, step);
  const int32_t nnz = tnz[ntiles];
  int32_t r0 = 0, r1 = 0, j0 = 0, j1 = 0;
  TileData<T, IX> d;
  if (t < end) {
    r0 = tiles[t]; r1 = tiles[t + 1]; j0 = tnz[t]; j1 = tnz[t + 1];
    d.load(ptr, col, val, r0, r1, j0, j1, nnz, tid);
  }
  // meta of the next tile: loaded one tile ahead, after the data prefetch,
  // so waiting for the prefetched columns never waits for these loads
  int32_t nr0 = 0, nr1 = 0, nj0 = 0, nj1 = 0;
  if (t + step < end) {
    nr0 = tiles[t + step]; nr1 = tiles[t + step + 1]; nj0 = tnz[t + step]; nj1 = tnz[t + step + 1];
  }
  // the epilogue's scalars, with the first tile already in flight (an
  // epilogue may reduce partials across the workgroup here)
  epi.prep();
  int slot = 0;  // this tile's index among the workgroup's tiles
  for (; t < end; t += step, ++slot) {
    const int nr = r1 - r0, cnt = j1 - j0;
    const int64_t tn = t + step, tnn = t + 2 * step;
    const bool more = tn < end, more2 = tnn < end;
    int32_t mr0 = 0, mr1 = 0, mj0 = 0, mj1 = 0;
    if (cnt > kTileNnz) {
      // one long row: strided partial sums, then a fixed tree
      T s0 = 0, s1 = 0;
      for (int k = tid; k < cnt; k += kBlock) {
        const T a = ld_stream(val + j0 + k);
        int c;
        if constexpr (kDelta) {
          // not reached (no row is a tile of its own in this format); kept
          // correct: the column is the prefix sum of the row's codes
          c = r0 - kDeltaBias;
          for (int q = 0; q <= k; ++q) c += (int)col[j0 + q];
        } else {
          c = (int)ld_stream(col + j0 + k) + (kRel ? r0 : 0);
        }
        s0 += a * x0[c];
        if (NX == 2) s1 += a * x1[c];
      }
      const int lane = tid & 63, w = tid >> 6;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, 64);
        if (NX == 2) s1 += __shfl_xor(s1, off, 64);
      }
      if (lane == 0) {
        slong[0][w] = (double)s0;
        if (NX == 2) slong[NX - 1][w] = (double)s1;
      }
      __syncthreads();
      if (tid == 0) {
        T a0 = 0, a1 = 0;
        for (int k = 0; k < kBlock / 64; ++k) {
          a0 += (T)slong[0][k];
          if (NX == 2) a1 += (T)slong[NX - 1][k];
        }
        if constexpr (kYT > 0) {
          if (slot < kYT) sy[slot][r0 & 3] = epi.value(a0);
          else epi.row(r0, a0, a1, acc);
        } else {
          epi.row(r0, a0, a1, acc);
        }
      }
      if (more) d.load(ptr, col, val, nr0, nr1, nj0, nj1, nnz, tid);
      if (more2) { mr0 = tiles[tnn]; mr1 = tiles[tnn + 1]; mj0 = tnz[tnn]; mj1 = tnz[tnn + 1]; }
      __syncthreads();
    } else {
      if (tid < nr) sptr[tid] = d.pr - d.a0of(j0);
      if (tid == 0) sptr[nr] = j1 - d.a0of(j0);
      d.stage(sv, sc, j0, j1, tid);
      __syncthreads();
      // in flight during the sums: the next tile's data, then the meta after it
      if (more) d.load(ptr, col, val, nr0, nr1, nj0, nj1, nnz, tid);
      if (more2) { mr0 = tiles[tnn]; mr1 = tiles[tnn + 1]; mj0 = tnz[tnn]; mj1 = tnz[tnn + 1]; }
      // more than one thread per row only for rows averaging > 8 entries
      // (delta-coded columns: always one per row, which the host checked is
      // the choice below for every tile)
      int tpr = 1;
      if constexpr (!kDelta)
        while (tpr < 64 && tpr * 2 * nr <= kBlock && tpr * 8 * nr < cnt) tpr *= 2;
      const int grp = tid / tpr, lane = tid % tpr;
      if (grp < nr) {
        const int s = sptr[grp], e = sptr[grp + 1];
        const int cb = kRel ? r0 + grp : kDelta ? r0 + grp - kDeltaBias : 0;
        int cd = cb;  // delta-coded: the running column
        T a0 = 0, a1 = 0;
#pragma unroll 2
        for (int k = s + lane; k < e; k += tpr) {
          const T v = sv[k];
          int c;
          if constexpr (kDelta) {
            cd += (int)sc[k];
            c = cd;
          } else {
            c = cb + (int)sc[k];
          }
          a0 = mul_add_rn(a0, v, x0[c]);
          if (NX == 2) a1 = mul_add_rn(a1, v, x1[c]);
        }
        for (int off = tpr >> 1; off > 0; off >>= 1) {
          a0 += __shfl_xor(a0, off, 64);
          if (NX == 2) a1 += __shfl_xor(a1, off, 64);
        }
        if (lane == 0) {
          if constexpr (kYT > 0) {
            if (slot < kYT) sy[slot][(r0 & 3) + grp] = epi.value(a0);
            else epi.row(r0 + grp, a0, a1, acc);
          } else {
            epi.row(r0 + grp, a0, a1, acc);
          }
        }
      }
      __syncthreads();
    }
    r0 = nr0; r1 = nr1; j0 = nj0; j1 = nj1;
    nr0 = mr0; nr1 = mr1; nj0 = mj0; nj1 = mj1;
  }
  if constexpr (NV > 0) block_store<NV>(acc, ra);
  if constexpr (HasProj<Epi>::value) {
    // Classical Gram-Schmidt projection fused behind the SpMV: the partial
    // dots (y, V_v) over this workgroup's own rows, which it has just
    // written (visible to the whole workgroup after the loop's last
    // barrier; still in this XCD's L2), so y is not read back from HBM and
    // no separate projection launch runs.  The four wavefronts split the
    // work by columns and by tiles: gc column groups of cpg = nv / gc
    // columns (gc = 1, 2 or 4, the fewest with cpg <= CP when possible),
    // each taken by 4 / gc wavefronts that split the workgroup's tiles
    // among them (every (4/gc)-th tile), CP columns per pass, with the next
    // tile's bounds loaded while the current one is summed.  So every
    // wavefront has work at every column count, and a tile costs one round
    // trip per pass.  Rows are read in 16-byte chunks (y, V and ldv 16-byte
    // aligned, checked by the host; a tile of <= 256 rows is at most NCH
    // chunks per lane), a tile's rows outside its aligned chunks one per
    // lane.  Per pass the sums of a column's wavefronts are added in a
    // fixed order through LDS and written to part[v * kMaxGrid + workgroup].
    static_assert(NV == 0 && NX == 1, "projection epilogue takes a plain SpMV");
    using VT = typename Chunk<T, 16 / sizeof(T)>::type;
    constexpr int W = 16 / sizeof(T);
    constexpr int kW = kBlock / 64;
    static_assert(kW == 4, "four wavefronts per workgroup");
    constexpr int CP = sizeof(T) == 4 ? 6 : 4;   // columns per pass (the most without spills)
    constexpr int NCH = kTileRows / (64 * W);     // aligned chunks per lane per tile
    __shared__ double sred[kW][CP];
    // the wavefront index is uniform across the wavefront: scalar registers
    // for the tile loop and its bounds
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    int64_t u0, uend, ustep;
    rbs.range(u0, uend, ustep);
    const int nv = epi.nv;
    const int64_t ldv = epi.ldv;
    const int gc = nv <= CP ? 1 : nv <= 2 * CP ? 2 : 4;
    const int gt = kW / gc, cpg = (nv + gc - 1) / gc;
    const int cbeg = (wv / gt) * cpg, cend = min(nv, cbeg + cpg);
    const int64_t ws = (int64_t)gt * ustep;
    // at least one pass: its column group 0 also writes the staged y out
    for (int p0 = 0; p0 == 0 || p0 < cpg; p0 += CP) {
      const int c0 = cbeg + p0;  // columns c0 + j < cend in this pass
      double pa[CP];
#pragma unroll
      for (int j = 0; j < CP; ++j) pa[j] = 0.0;
      const T* vb = epi.V + (int64_t)c0 * ldv;
      int64_t u = u0 + (int64_t)(wv % gt) * ustep;
      int ti = wv % gt;  // the tile's index among the workgroup's tiles
      int64_t q0 = 0, q1 = 0;
      if (u < uend) { q0 = tiles[u]; q1 = tiles[u + 1]; }
      // one tile: y from LDS (a staged tile, kStaged) or from memory; the
      // staged tiles come first in every wavefront's order, so two loops
      // keep each load in one address space
      auto tile = [&](auto staged_tag) {
        constexpr bool kStaged = decltype(staged_tag)::value;
        int64_t n0 = 0, n1 = 0;
        if (u + ws < uend) { n0 = tiles[u + ws]; n1 = tiles[u + ws + 1]; }
        const int64_t a = (q0 + W - 1) / W * W, b = q1 / W * W;
        const int64_t base = q0 & ~(int64_t)3;
        const T* ys = nullptr;  // row r of a staged tile at ys[r - base]
        if constexpr (kStaged) {
          ys = sy[ti];
          if (p0 == 0 && wv < gt)
            for (int64_t j = q0 + lane; j < q1; j += 64) epi.y[j] = ys[j - base];
        }
        {
          VT yv[NCH], xv[NCH][CP];
#pragma unroll
          for (int h = 0; h < NCH; ++h) {
            const int64_t r = a + (int64_t)W * (lane + 64 * h);
            if (r + W <= b) {
              if constexpr (kStaged) yv[h] = *reinterpret_cast<const VT*>(ys + (r - base));
              else yv[h] = *reinterpret_cast<const VT*>(epi.y + r);
#pragma unroll
              for (int j = 0; j < CP; ++j)
                if (c0 + j < cend) xv[h][j] = *reinterpret_cast<const VT*>(vb + (int64_t)j * ldv + r);
            }
          }
#pragma unroll
          for (int h = 0; h < NCH; ++h) {
            const int64_t r = a + (int64_t)W * (lane + 64 * h);
            if (r + W <= b) {
#pragma unroll
              for (int j = 0; j < CP; ++j)
                if (c0 + j < cend) {
                  double t = 0.0;
#pragma unroll
                  for (int e = 0; e < W; ++e) t += (double)yv[h][e] * (double)xv[h][j][e];
                  pa[j] += t;
                }
            }
          }
        }
        // rows before the first and after the last aligned chunk (< 2W; none
        // on tiles with chunk-aligned bounds), after the chunks
        const int64_t hn = (a < q1 ? a : q1) - q0;
        const int64_t tb = b > a ? b : a;
        int64_t row = -1;
        if (lane < hn) row = q0 + lane;
        else if (lane >= W && lane - W < q1 - tb) row = tb + (lane - W);
        if (row >= 0) {
          double ye;
          if constexpr (kStaged) ye = (double)ys[row - base];
          else ye = (double)epi.y[row];
#pragma unroll
          for (int j = 0; j < CP; ++j)
            if (c0 + j < cend) pa[j] += ye * (double)vb[(int64_t)j * ldv + row];
        }
        q0 = n0; q1 = n1;
      };
      if constexpr (kYT > 0)
        for (; u < uend && ti < kYT; u += ws, ti += gt) tile(std::true_type{});
      for (; u < uend; u += ws, ti += gt) tile(std::false_type{});
#pragma unroll
      for (int j = 0; j < CP; ++j) {
        const double v = wave_sum(pa[j]);
        if (lane == 0) sred[wv][j] = v;
      }
      __syncthreads();
      if (tid < gc * CP) {
        const int g = tid / CP, j = tid % CP, c = g * cpg + p0 + j;
        if (p0 + j < cpg && c < nv) {
          const double* q = &sred[g * gt][j];
          double t;
          if (gt == 4) t = (q[0] + q[CP]) + (q[2 * CP] + q[3 * CP]);
          else if (gt == 2) t = q[0] + q[CP];
          else t = q[0];
          ra.part[(size_t)c * kMaxGrid + blockIdx.x] = t;
        }
      }
      __syncthreads();
    }
  }
}

// Wavefront CSR-stream: the scheme of k_spmv_stream (values and column
// indices staged in LDS, one row per lane for short rows) with every wavefront
// working alone on tiles of at most kWaveRows rows and kWaveNnz entries
// (8 per lane), so no workgroup barrier sits in the tile loop and one slow
// gather stalls only its own wavefront.  A wavefront's LDS instructions
// execute in order, so wavefront-scope fences (which keep the compiler from
// moving LDS accesses across them) order its staging writes before its row
// sums and its sums before the next tile's writes.
constexpr int kWaveWin = 512;
constexpr int kWaveNnz = kWaveWin - 3;
constexpr int kWaveRows = 64;

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
using WaveTile = QuadTile<T, 64, kWaveWin>;

template <class T, int NX, int NV, class Epi>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kStreamWaves<T, NX, NV>))) void k_spmv_wave(
    int64_t n, int64_t ntiles, const int32_t* __restrict__ tiles, const int32_t* __restrict__ tnz,
    const int32_t* __restrict__ ptr, const int32_t* __restrict__ col, const T* __restrict__ val,
    const T* __restrict__ x0, const T* __restrict__ x1, Epi epi, Gate g, RedArgs ra, StatArgs st) {
  static_assert(!HasProj<Epi>::value, "the projection epilogue is a workgroup-tile kernel feature");
  if (g.off()) return;
  st.count();
  epi.prep();
  constexpr int kW = kBlock / 64;
  __shared__ __align__(16) T sv_all[kW][kWaveWin];
  __shared__ __align__(16) int32_t sc_all[kW][kWaveWin];
  __shared__ int32_t sptr_all[kW][kWaveRows + 1];
  double acc[NV > 0 ? NV : 1];
#pragma unroll
  for (int i = 0; i < (NV > 0 ? NV : 1); ++i) acc[i] = 0.0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T* sv = sv_all[w];
  int32_t* sc = sc_all[w];
  int32_t* sptr = sptr_all[w];
  RowBlocks rbs{ntiles};
  int64_t t, end, step;
  rbs.range(t, end, step, w, kW);
  const int32_t nnz = tnz[ntiles];
  int32_t r0 = 0, r1 = 0, j0 = 0, j1 = 0;
  WaveTile<T> d;
  if (t < end) {
    r0 = tiles[t]; r1 = tiles[t + 1]; j0 = tnz[t]; j1 = tnz[t + 1];
    d.load(ptr, col, val, r0, r1, j0, j1, nnz, lane);
  }
  int32_t nr0 = 0, nr1 = 0, nj0 = 0, nj1 = 0;
  if (t + step < end) {
    nr0 = tiles[t + step]; nr1 = tiles[t + step + 1]; nj0 = tnz[t + step]; nj1 = tnz[t + step + 1];
  }
  for (; t < end; t += step) {
    const int nr = r1 - r0, cnt = j1 - j0;
    const int64_t tn = t + step, tnn = t + 2 * step;
    const bool more = tn < end, more2 = tnn < end;
    int32_t mr0 = 0, mr1 = 0, mj0 = 0, mj1 = 0;
    if (cnt > kWaveNnz) {
      // one long row: lane-strided partial sums, then a fixed shuffle tree
      T s0 = 0, s1 = 0;
      for (int k = lane; k < cnt; k += 64) {
        const T a = ld_stream(val + j0 + k);
        const int c = ld_stream(col + j0 + k);
        s0 += a * x0[c];
        if (NX == 2) s1 += a * x1[c];
      }
      if (more) d.load(ptr, col, val, nr0, nr1, nj0, nj1, nnz, lane);
      if (more2) { mr0 = tiles[tnn]; mr1 = tiles[tnn + 1]; mj0 = tnz[tnn]; mj1 = tnz[tnn + 1]; }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, 64);
        if (NX == 2) s1 += __shfl_xor(s1, off, 64);
      }
      if (lane == 0) epi.row(r0, s0, s1, acc);
    } else {
      if (lane < nr) sptr[lane] = d.pr - (j0 & ~3);
      if (lane == 0) sptr[nr] = j1 - (j0 & ~3);
      d.stage(sv, sc, j0, j1, lane);
      wave_sync();
      // in flight during the sums: the next tile's data, then the bounds after it
      if (more) d.load(ptr, col, val, nr0, nr1, nj0, nj1, nnz, lane);
      if (more2) { mr0 = tiles[tnn]; mr1 = tiles[tnn + 1]; mj0 = tnz[tnn]; mj1 = tnz[tnn + 1]; }
      // more than one lane per row only for rows averaging > 8 entries
      int tpr = 1;
      while (tpr < 64 && tpr * 2 * nr <= 64 && tpr * 8 * nr < cnt) tpr *= 2;
      const int grp = lane / tpr, sub = lane % tpr;
      T a0 = 0, a1 = 0;
      if (grp < nr) {
        const int s = sptr[grp], e = sptr[grp + 1];
#pragma unroll 4
        for (int k = s + sub; k < e; k += tpr) {
          const T v = sv[k];
          const int c = sc[k];
          a0 = mul_add_rn(a0, v, x0[c]);
          if (NX == 2) a1 = mul_add_rn(a1, v, x1[c]);
        }
      }
      for (int off = tpr >> 1; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, 64);
        if (NX == 2) a1 += __shfl_xor(a1, off, 64);
      }
      if (grp < nr && sub == 0) epi.row(r0 + grp, a0, a1, acc);
      wave_sync();
    }
    r0 = nr0; r1 = nr1; j0 = nj0; j1 = nj1;
    nr0 = mr0; nr1 = mr1; nj0 = mj0; nj1 = mj1;
  }
  if constexpr (NV > 0) block_store<NV>(acc, ra);
}

// out[j] = alpha * u[row(j)] * v[col(j)] over the pattern (SDDMM).
template <int L>
__global__ __launch_bounds__(kBlock) void k_sddmm(int64_t n, const int32_t* __restrict__ ptr,
                                                  const int32_t* __restrict__ col, double alpha,
                                                  const double* __restrict__ u,
                                                  const double* __restrict__ v, double* __restrict__ out,
                                                  StatArgs st) {
  st.count();
  constexpr int RPB = kBlock / L;
  __shared__ int32_t sptr[RPB + 1];
  const int grp = threadIdx.x / L, lane = threadIdx.x % L;
  RowBlocks rbs{(n + RPB - 1) / RPB};
  int64_t rb, end, step;
  rbs.range(rb, end, step);
  for (; rb < end; rb += step) {
    const int64_t row0 = rb * RPB;
    __syncthreads();
    if (threadIdx.x <= RPB) {
      int64_t r = row0 + threadIdx.x;
      sptr[threadIdx.x] = ptr[r < n ? r : n];
    }
    __syncthreads();
    const int64_t row = row0 + grp;
    if (row < n) {
      const double ur = alpha * u[row];
      const int s = sptr[grp], e = sptr[grp + 1];
      for (int j = s + lane; j < e; j += L) out[j] = ur * v[col[j]];
    }
  }
}

// ---------------------------------------------------- merge-path SpMV --
// The merge path of the row ends (ptr[1..n]) and the entries (0..nnz) is cut
// into segments of kMergeItems items (Merrill & Garland, "Merge-based
// parallel sparse matrix-vector multiplication", SC16); segment s starts at
// path point (segr[s], segj[s]) (computed on the host), so it holds at most
// kMergeItems entries, [segj[s], segj[s+1]).  A row belongs to the segment in
// which it ends.  One pass over the entries (k_spmv_merge) leaves, per
// segment, up to two partial sums: carry[s], the part of the row still open
// at the segment's end, and head[s], the part of the row continued from
// earlier segments that ends here.  Every other row of the segment starts
// and ends in it and passes through the epilogue in the same pass.
// k_merge_fix then forms each continued row's total as (its earlier
// segments' carries, added in segment order) + head, and runs its
// epilogue - so every row total forms in a fixed order.
constexpr int kMergeItems = 512;
constexpr int kMergeUnroll = kMergeItems / 64;
// workgroups of k_merge_fix (its reduction partials follow the main
// kernel's in the partial buffer)
constexpr int kMergeFixGrid = 64;

// Register budget of k_spmv_merge: two segments' entries live at once (the
// current one's and the next one's loads).
template <class T, int NX>
constexpr int kMergeWaves = sizeof(T) == 4 ? 6 : (NX == 2 ? 4 : 5);

// One pass of the merge path.  Each wavefront takes its segments in order
// and keeps the next segment's bounds and entries in flight while it
// gathers, stages and sums the current one:
//   1. gathers for the current entries; loads of the next entries, the
//      bounds two segments ahead and the current rows' pointers;
//   2. products v * x (rounded, no FMA) of all the segment's entries to LDS;
//   3. carry[s]: lane l adds the open row's products l, l+64, ... in order,
//      then a fixed shuffle tree;
//   4. the segment's rows from LDS, tpr lanes per row (a fixed tree); the
//      continued row's sum goes to head[s] instead of the epilogue.
// carry and head hold NX values per segment: carry[v * nseg + s],
// head[v * nseg + s] = carry[(NX + v) * nseg + s].
template <class T, int NX, int NV, class Epi>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kMergeWaves<T, NX>))) void k_spmv_merge(
    int64_t n, int64_t nseg, const int32_t* __restrict__ segr, const int32_t* __restrict__ segj,
    const int32_t* __restrict__ ptr, const int32_t* __restrict__ col, const T* __restrict__ val,
    const T* __restrict__ x0, const T* __restrict__ x1, T* __restrict__ carry, Epi epi, Gate g, RedArgs ra,
    StatArgs st) {
  if (g.off()) return;
  st.count();
  epi.prep();
  constexpr int kW = kBlock / 64;
  constexpr int U = kMergeUnroll;
  __shared__ T prod[NX][kW][kMergeItems];
  double acc[NV > 0 ? NV : 1];
#pragma unroll
  for (int i = 0; i < (NV > 0 ? NV : 1); ++i) acc[i] = 0.0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T* p0 = prod[0][w];
  T* p1 = prod[NX - 1][w];
  T* head = carry + (int64_t)NX * nseg;
  RowBlocks rbs{nseg};
  int64_t s, end, step;
  rbs.range(s, end, step, w, kW);
  // current segment: bounds and entries; next segment: bounds and entries
  // in flight; the one after: bounds
  int32_t r0 = 0, r1 = 0, js = 0, je = 0, nr0 = 0, nr1 = 0, njs = 0, nje = 0;
  T v[U], vn[U];
  int32_t c[U], cn[U];
#pragma unroll
  for (int q = 0; q < U; ++q) { v[q] = vn[q] = T(0); c[q] = cn[q] = 0; }
  if (s < end) {
    r0 = segr[s]; r1 = segr[s + 1]; js = segj[s]; je = segj[s + 1];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int32_t k = js + lane + 64 * q;
      if (k < je) { v[q] = ld_stream(val + k); c[q] = ld_stream(col + k); }
    }
  }
  if (s + step < end) { nr0 = segr[s + step]; nr1 = segr[s + step + 1]; njs = segj[s + step]; nje = segj[s + step + 1]; }
  for (; s < end; s += step) {
    const bool more = s + step < end, more2 = s + 2 * step < end;
    const int cnt = je - js;
    // 1. gathers; then the next segment's entries and the bounds after it
    T gx0[U], gx1[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      gx0[q] = gx1[q] = T(0);
      if (lane + 64 * q < cnt) {
        gx0[q] = x0[c[q]];
        if (NX == 2) gx1[q] = x1[c[q]];
      }
    }
    if (more) {
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const int32_t k = njs + lane + 64 * q;
        if (k < nje) { vn[q] = ld_stream(val + k); cn[q] = ld_stream(col + k); }
      }
    }
    int32_t mr0 = 0, mr1 = 0, mjs = 0, mje = 0;
    if (more2) {
      mr0 = segr[s + 2 * step]; mr1 = segr[s + 2 * step + 1]; mjs = segj[s + 2 * step]; mje = segj[s + 2 * step + 1];
    }
    // the row pointers of the segment's first 64 rows (row r0 + lane), and
    // the open row's start
    const int nr = r1 - r0;
    const int32_t prow = ptr[r0 + (lane < nr ? lane : nr)];
    const int32_t pr1 = r1 < n ? ptr[r1] : je;
    // 2. products to LDS
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int k = lane + 64 * q;
      if (k < cnt) {
        p0[k] = v[q] * gx0[q];
        if (NX == 2) p1[k] = v[q] * gx1[q];
      }
    }
    wave_sync();
    // 3. the open row's part: carry[s]
    {
      T a0 = 0, a1 = 0;
      if (r1 < n) {
        const int a = (pr1 > js ? pr1 : js) - js;
        for (int k = a + lane; k < cnt; k += 64) {
          a0 += p0[k];
          if (NX == 2) a1 += p1[k];
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, 64);
        if (NX == 2) a1 += __shfl_xor(a1, off, 64);
      }
      if (lane == 0) {
        carry[s] = a0;
        if (NX == 2) carry[nseg + s] = a1;
      }
    }
    // 4. the segment's rows r0 .. r1 - 1
    const int ecnt = (pr1 < je ? pr1 : je) - js;  // entries of rows that end here
    int tpr = 1;
    while (tpr < 64 && tpr * 2 * nr <= 64 && tpr * 8 * nr < ecnt) tpr *= 2;
    const int sub = lane % tpr;
    for (int base = 0; base < nr; base += 64 / tpr) {
      const int grp = base + lane / tpr;
      const bool live = grp < nr;
      int32_t pa = __shfl(prow, grp < 63 ? grp : 63, 64), pb = __shfl(prow, grp < 62 ? grp + 1 : 63, 64);
      if (live && grp >= 62) {
        pa = ptr[r0 + grp];
        pb = ptr[r0 + grp + 1];
      }
      T a0 = 0, a1 = 0;
      if (live) {
        const int sa = (pa > js ? pa : js) - js, sb = pb - js;
        for (int k = sa + sub; k < sb; k += tpr) {
          a0 += p0[k];
          if (NX == 2) a1 += p1[k];
        }
      }
      for (int off = tpr >> 1; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, 64);
        if (NX == 2) a1 += __shfl_xor(a1, off, 64);
      }
      if (live && sub == 0) {
        if (grp == 0 && pa < js) {  // continued from earlier segments: k_merge_fix finishes it
          head[s] = a0;
          if (NX == 2) head[nseg + s] = a1;
        } else {
          epi.row(r0 + grp, a0, a1, acc);
        }
      }
    }
    wave_sync();
    r0 = nr0; r1 = nr1; js = njs; je = nje;
    nr0 = mr0; nr1 = mr1; njs = mjs; nje = mje;
#pragma unroll
    for (int q = 0; q < U; ++q) { v[q] = vn[q]; c[q] = cn[q]; }
  }
  if constexpr (NV > 0) block_store<NV>(acc, ra);
}

// The rows continued from earlier segments that end in segment s: total =
// (carries of the segments f .. s-1 whose open row it was, in order) +
// head[s], then the epilogue; the reduction partials go to this kernel's own
// slots (ra points past the main kernel's).
template <class T, int NX, int NV, class Epi>
__global__ __launch_bounds__(kBlock) void k_merge_fix(int64_t n, int64_t nseg, const int32_t* __restrict__ segr,
                                                      const int32_t* __restrict__ segj,
                                                      const int32_t* __restrict__ ptr,
                                                      const T* __restrict__ carry, Epi epi, Gate g, RedArgs ra) {
  if (g.off()) return;
  epi.prep();
  double acc[NV > 0 ? NV : 1];
#pragma unroll
  for (int i = 0; i < (NV > 0 ? NV : 1); ++i) acc[i] = 0.0;
  const T* head = carry + (int64_t)NX * nseg;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += stride) {
    const int32_t r0 = segr[s], r1 = segr[s + 1];
    if (r0 < r1 && ptr[r0] < segj[s]) {
      int64_t f = s;
      while (f > 0 && segr[f] == r0) --f;
      T c0 = 0, c1 = 0;
      for (int64_t q = f; q < s; ++q) {
        c0 += carry[q];
        if (NX == 2) c1 += carry[nseg + q];
      }
      epi.row(r0, c0 + head[s], NX == 2 ? c1 + head[nseg + s] : T(0), acc);
    }
  }
  if constexpr (NV > 0) block_store<NV>(acc, ra);
}

// ---------------------------------------------------- Gram-Schmidt ------
// 16-byte vectors of T (float4 / double2) for the column-streaming kernels:
// with VEC, thread i reads elements [i*N, i*N + N) of every column with one
// 16-byte load (the columns and w must be 16-byte aligned and ldv a multiple
// of N); the n % N tail elements are taken by the first threads.
template <class T>
struct Vec16 {
  static constexpr int N = 16 / sizeof(T);
  typedef T type __attribute__((ext_vector_type(16 / sizeof(T))));
};

// h[i] = (w, V_{i0+i}) for i < nv (nv <= G): one pass over w and nv columns.
template <class T, int G, bool VEC>
__global__ __launch_bounds__(kBlock) void k_proj(int64_t n, const T* __restrict__ V, int64_t ldv, int nv,
                                                 const T* __restrict__ w, Gate g, RedArgs ra, StatArgs st) {
  if (g.off()) return;
  st.count();
  double acc[G];
#pragma unroll
  for (int v = 0; v < G; ++v) acc[v] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if constexpr (VEC) {
    using VT = typename Vec16<T>::type;
    constexpr int N = Vec16<T>::N;
    const int64_t nvec = n / N;
    const VT* wv = reinterpret_cast<const VT*>(w);
    for (int64_t i = gid; i < nvec; i += stride) {
      const VT a = wv[i];
#pragma unroll
      for (int v = 0; v < G; ++v)
        if (v < nv) {
          const VT b = reinterpret_cast<const VT*>(V + (int64_t)v * ldv)[i];
          double t = 0.0;
#pragma unroll
          for (int e = 0; e < N; ++e) t += (double)a[e] * (double)b[e];
          acc[v] += t;
        }
    }
    for (int64_t i = nvec * N + gid; i < n; i += stride) {
      const double wi = (double)w[i];
#pragma unroll
      for (int v = 0; v < G; ++v)
        if (v < nv) acc[v] += wi * (double)V[(int64_t)v * ldv + i];
    }
  } else {
    int64_t i = gid;
    for (; i + stride < n; i += 2 * stride) {
      const double wa = (double)w[i], wb = (double)w[i + stride];
#pragma unroll
      for (int v = 0; v < G; ++v)
        if (v < nv) acc[v] += wa * (double)V[(int64_t)v * ldv + i] + wb * (double)V[(int64_t)v * ldv + i + stride];
    }
    for (; i < n; i += stride) {
      const double wi = (double)w[i];
#pragma unroll
      for (int v = 0; v < G; ++v)
        if (v < nv) acc[v] += wi * (double)V[(int64_t)v * ldv + i];
    }
  }
  block_store<G>(acc, ra);
}
// Finisher of k_proj: h column entries.  With vn set, column v of the basis
// is stored unnormalised with norm vn[v], and h_v = (w, V_v) / vn[v].
struct FinProj {
  double* hout;
  int ldh, nv;
  const double* vn;
  // in chunks of 8: the chunk's norms are loaded before its stores (hout and
  // vn do not overlap), one memory round trip per chunk
  __device__ void fin(const double* s) const {
    for (int v0 = 0; v0 < nv; v0 += 8) {
      double d[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) d[j] = (vn && v0 + j < nv) ? vn[v0 + j] : 1.0;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (v0 + j < nv) hout[(int64_t)(v0 + j) * ldh] = vn ? s[v0 + j] / d[j] : s[v0 + j];
    }
  }
};

// out = w - sum_{i<nv} h_i V_i (classical GS update; out may be w), then
// ||out||^2 -> fin.  With vn set, column i of V is stored unnormalised with
// norm vn[i] and enters as h_i / vn[i].
template <class T, bool VEC>
__global__ __launch_bounds__(kBlock) void k_cgs_update(int64_t n, const T* __restrict__ V, int64_t ldv,
                                                       int nv, const T* w, T* out, const double* hcol,
                                                       int ldh, const double* vn, Gate g, RedArgs ra,
                                                       StatArgs st) {
  if (g.off()) return;
  st.count();
  extern __shared__ double sh[];
  for (int v = threadIdx.x; v < nv; v += kBlock) sh[v] = vn ? hcol[(int64_t)v * ldh] / vn[v] : hcol[(int64_t)v * ldh];
  __syncthreads();
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int64_t i0 = gid;
  if constexpr (VEC) {
    using VT = typename Vec16<T>::type;
    constexpr int N = Vec16<T>::N;
    const int64_t nvec = n / N;
    const VT* wv = reinterpret_cast<const VT*>(w);
    VT* ov = reinterpret_cast<VT*>(out);
    for (int64_t i = gid; i < nvec; i += stride) {
      const VT a = wv[i];
      double t[N];
#pragma unroll
      for (int e = 0; e < N; ++e) t[e] = (double)a[e];
#pragma unroll 4
      for (int v = 0; v < nv; ++v) {
        const VT b = reinterpret_cast<const VT*>(V + (int64_t)v * ldv)[i];
        const double h = sh[v];
#pragma unroll
        for (int e = 0; e < N; ++e) t[e] -= h * (double)b[e];
      }
      VT o;
#pragma unroll
      for (int e = 0; e < N; ++e) {
        o[e] = (T)t[e];
        acc[0] += (double)o[e] * (double)o[e];
      }
      ov[i] = o;
    }
    i0 = nvec * N + gid;
  }
  for (int64_t i = i0; i < n; i += stride) {
    double t = (double)w[i];
#pragma unroll 8
    for (int v = 0; v < nv; ++v) t -= sh[v] * (double)V[(int64_t)v * ldv + i];
    const T tw = (T)t;
    out[i] = tw;
    acc[0] += (double)tw * (double)tw;
  }
  block_store<1>(acc, ra);
}

// x += sum_{i<size} y_i B_i   (or only the last term with KS_COMPAT_LAST_TERM
// when the cycle stopped inside); y and size come from the Ctl (k_hye).
// Tx: type of x, Tb: type of the basis B.
// With c->vn set (and use_vn), B holds the unnormalised basis: y_i / vn[i].
// With xw set, the new x is also written there widened to FP64 (an FP32
// solve's output for the level above it).  While c->xzero is set (the first
// cycle from x0 = 0) x is taken as 0.0 without being read, so the zero is
// never written either.
template <class Tx, class Tb>
__global__ __launch_bounds__(kBlock) void k_assemble(int64_t n, const Tb* __restrict__ B, int64_t ldb,
                                                     Tx* __restrict__ x, double* __restrict__ xw, const Ctl* c,
                                                     int last_term, int use_vn, Gate g, StatArgs st) {
  if (g.off()) return;
  st.count();
  extern __shared__ double sy[];
  __shared__ int s_sz, s_lt, s_xz;
  if (threadIdx.x == 0) {
    s_sz = c->stop ? c->ksz : c->R - 1;
    s_lt = last_term && c->stop;
    s_xz = c->xzero;
  }
  __syncthreads();
  const int sz = s_sz;
  for (int v = threadIdx.x; v < sz; v += kBlock) sy[v] = use_vn ? c->y[v] / c->vn[v] : c->y[v];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  if (s_lt) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
      double t = sz > 0 ? sy[sz - 1] * (double)B[(int64_t)(sz - 1) * ldb + i] : 0.0;
      const Tx xi = (Tx)((s_xz ? 0.0 : (double)x[i]) + (double)(Tb)t);
      x[i] = xi;
      if (xw) xw[i] = (double)xi;
    }
  } else {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
      double t = 0.0;
#pragma unroll 8
      for (int v = 0; v < sz; ++v) t += sy[v] * (double)B[(int64_t)v * ldb + i];
      const Tx xi = (Tx)((s_xz ? 0.0 : (double)x[i]) + (double)(Tb)t);
      x[i] = xi;
      if (xw) xw[i] = (double)xi;
    }
  }
}

// GCR block update (blas::Gcr_sp_1 + Vec_add, blas.cpp:526-543):
// beta_i = -(Av, q_i)/qq_i (hq = (Av, q_i) from k_proj), then
// P_{k+1} = z + sum_i beta_i P_i and Q_{k+1} = Av + sum_i beta_i Q_i with the
// sums accumulated in i order from 0 (the reference's Scalar_axy sequence).
template <class T>
__global__ __launch_bounds__(kBlock) void k_gcr_update(int64_t n, const T* __restrict__ P, const T* __restrict__ Q,
                                                       int64_t ldv, int nv, const T* __restrict__ z,
                                                       const T* __restrict__ Av, const double* hq,
                                                       const double* qq, T* __restrict__ Pn, T* __restrict__ Qn,
                                                       Gate g, StatArgs st) {
  if (g.off()) return;
  st.count();
  extern __shared__ double sb[];
  for (int v = threadIdx.x; v < nv; v += kBlock) sb[v] = -hq[v] / qq[v];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    double tp = 0.0, tq = 0.0;
    for (int v = 0; v < nv; ++v) {
      tp = sb[v] * (double)P[(int64_t)v * ldv + i] + tp;
      tq = sb[v] * (double)Q[(int64_t)v * ldv + i] + tq;
    }
    Pn[i] = (T)((double)z[i] + (double)(T)tp);
    Qn[i] = (T)((double)Av[i] + (double)(T)tq);
  }
}

// ------------------------------------------------------------ GMRES ----
// Back substitution y = H(0:sz,0:sz)^{-1} e (blas.cpp:303-314).
__global__ void k_hye(Ctl* c, Gate g) {
  if (g.off()) return;
  const int sz = c->stop ? c->ksz : c->R - 1;
  const int ld = c->ld;
  for (int i = 0; i < sz; ++i) c->y[i] = 0.0;
  for (int i = sz - 1; i >= 0; --i) {
    double t = 0.0;
    for (int j = i + 1; j < sz; ++j) t += c->y[j] * c->H[i * ld + j];
    c->y[i] = (c->e[i] - t) / c->H[i * ld + i];
  }
}

// Marks the solve done once the stopping cycle has been assembled.
__global__ void k_mark_done(Ctl* c, Gate g) {
  if (g.off()) return;
  if (c->stop) c->done = 1;
}

// Givens rotations for column k (gmres.cpp:384-401), then the stop test of
// step k+1 when it lies in this cycle (gmres.cpp:278-324).  Runs in the
// finisher of the Gram-Schmidt norm, right after h_{k+1,k} is known.
// GMRES step k after the Gram-Schmidt norm h = ||w||: H[k+1][k] = h, the
// Givens rotations of column k (gmres.cpp:384-399), the residual estimate
// |e_{k+1}| / ||b|| and the stop test (gmres.cpp:278, the `count+1 >=
// maxloop` rule at :286).  Runs on one thread, so it is written as loads,
// then arithmetic, then stores: every control field and array entry it needs
// is read before its first store (H, cs, sn, e, vn and the control block do
// not overlap), and the column is rotated in chunks of 8 whose entries are
// loaded together - a few memory round trips per step instead of one per
// field and per rotation.
__device__ inline void gm_norm_step(Ctl* c, int k, double h) {
  double* __restrict__ H = c->H;
  double* __restrict__ cs = c->cs;
  double* __restrict__ sn = c->sn;
  double* __restrict__ e = c->e;
  double* __restrict__ vn = c->vn;
  double* __restrict__ vr = c->vr;
  const int ld = c->ld, R = c->R;
  const long long count = c->count, steps = c->steps, tsteps = c->tsteps, maxloop = c->maxloop;
  const long long nhist = c->nhist, hist_cap = c->hist_cap;
  double* hist = c->hist;
  int* host_stop = c->host_stop;
  const double bnorm = c->bnorm, eps = c->eps;
  const double ek = e[k];
  double t1 = H[k];  // H[0][k], then the partly rotated H[i][k]
  for (int i0 = 0; i0 < k; i0 += 8) {
    double t2[8], cv[8], sv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (i0 + j < k) {
        t2[j] = H[(i0 + j + 1) * ld + k];
        cv[j] = cs[i0 + j];
        sv[j] = sn[i0 + j];
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (i0 + j < k) {
        H[(i0 + j) * ld + k] = cv[j] * t1 - sv[j] * t2[j];
        t1 = sv[j] * t1 + cv[j] * t2[j];
      }
    }
  }
  // t1 is now the rotated H[k][k]; H[k+1][k] = h is annihilated
  const double hk = t1;
  const double alpha = sqrt(hk * hk + h * h);
  const double ck = hk / alpha, sk = -h / alpha;
  const double ek1 = sk * ek;
  cs[k] = ck;
  sn[k] = sk;
  e[k + 1] = ek1;
  e[k] = ck * ek;
  H[k * ld + k] = ck * hk - sk * h;
  H[(k + 1) * ld + k] = 0.0;
  vn[k + 1] = h;
  vr[k + 1] = 1.0 / h;
  c->vscale = h;
  c->count = count + 1;
  c->steps = steps + 1;
  c->tsteps = tsteps + 1;
  if (k + 1 < R - 1) {
    const double err = fabs(ek1) / bnorm;
    c->err = err;
    if (hist && nhist < hist_cap) {
      hist[2 * nhist] = (double)(count + 2);
      hist[2 * nhist + 1] = err;
    }
    c->nhist = nhist + 1;
    int st = -1;
    if (count + 2 >= maxloop) st = 2;
    else if (err <= eps) st = 0;
    else if (!finite(err)) st = 3;
    if (st >= 0) {
      c->ksz = k + 1;
      c->stop = 1;
      c->status = st;
      if (host_stop) *(volatile int*)host_stop = 1;
    }
  }
}

// ------------------------------------------------------ k-skip Gram -----
// CG: sums[0] = (r,r); delta_i = (r,Ar_i) i<2K; eta_i = (r,Ap_i) i<2K+1;
// zeta_i = (p,Ap_i) i<2K+2.  BiCG (shadow rs, ps): sums[0] = (rs,r),
// theta_i = (rs,Ar_i) i<2K, eta_i = (rs,Ap_i), rho_i = (ps,Ar_i) i<2K+1,
// phi_i = (ps,Ap_i) i<2K+2.  Ar/Ap are column-major blocks with stride ld.
// K is a template parameter so the accumulators stay in registers.
template <int K, bool BICG>
struct OpGram {
  static constexpr int NV = BICG ? 8 * K + 5 : 6 * K + 4;
  int64_t ld;
  const double *r, *p, *rs, *ps, *Ar, *Ap;
  Ctl* c;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double (&acc)[NV]) {
    if constexpr (!BICG) {
      const double ri = r[i], pi = p[i];
      acc[0] += ri * ri;
#pragma unroll
      for (int v = 0; v < 2 * K; ++v) acc[1 + v] += ri * Ar[v * ld + i];
#pragma unroll
      for (int v = 0; v < 2 * K + 2; ++v) {
        const double a = Ap[v * ld + i];
        if (v < 2 * K + 1) acc[1 + 2 * K + v] += ri * a;
        acc[1 + 4 * K + 1 + v] += pi * a;
      }
    } else {
      const double rsi = rs[i], psi = ps[i];
      acc[0] += rsi * r[i];
#pragma unroll
      for (int v = 0; v < 2 * K + 1; ++v) {
        const double a = Ar[v * ld + i];
        if (v < 2 * K) acc[1 + v] += rsi * a;
        acc[1 + 2 * K + (2 * K + 1) + v] += psi * a;
      }
#pragma unroll
      for (int v = 0; v < 2 * K + 2; ++v) {
        const double a = Ap[v * ld + i];
        if (v < 2 * K + 1) acc[1 + 2 * K + v] += rsi * a;
        acc[1 + 2 * K + 2 * (2 * K + 1) + v] += psi * a;
      }
    }
  }
  __device__ void fin(const double* s) {
    c->gamma = s[0];
    int o = 1;
    for (int v = 0; v < 2 * K; ++v) c->delta[v] = s[o + v];  // theta for BiCG
    o += 2 * K;
    for (int v = 0; v < 2 * K + 1; ++v) c->eta[v] = s[o + v];
    o += 2 * K + 1;
    if constexpr (BICG) {
      for (int v = 0; v < 2 * K + 1; ++v) c->rho[v] = s[o + v];
      o += 2 * K + 1;
      for (int v = 0; v < 2 * K + 2; ++v) c->phi[v] = s[o + v];
    } else {
      for (int v = 0; v < 2 * K + 2; ++v) c->zeta[v] = s[o + v];
    }
  }
};

// ------------------------------------------------------- transpose ------
__global__ void k_col_count(int64_t nnz, const int32_t* __restrict__ col, int32_t* cnt) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz; j += stride)
    atomicAdd(cnt + col[j], 1);
}

// Exclusive scan, phase 1: per-block sums of `in` (block of 1024 elements).
__global__ void k_scan_blocks(int64_t n, const int32_t* __restrict__ in, int64_t* bsum) {
  __shared__ int64_t s[kBlock];
  const int64_t base = (int64_t)blockIdx.x * 1024;
  int64_t t = 0;
  for (int q = 0; q < 4; ++q) {
    int64_t i = base + threadIdx.x * 4 + q;
    if (i < n) t += in[i];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = s[0];
}
// Phase 2: serial exclusive scan of the block sums (few thousand values).
__global__ void k_scan_serial(int64_t nb, int64_t* bsum) {
  int64_t acc = 0;
  for (int64_t b = 0; b < nb; ++b) {
    int64_t v = bsum[b];
    bsum[b] = acc;
    acc += v;
  }
  bsum[nb] = acc;
}
// Phase 3: out[i] = offset + exclusive prefix within the block; out[n] = total.
__global__ void k_scan_apply(int64_t n, const int32_t* __restrict__ in, const int64_t* bsum,
                             int32_t* out) {
  __shared__ int64_t s[kBlock];
  const int64_t base = (int64_t)blockIdx.x * 1024;
  int64_t v[4];
  int64_t t = 0;
  for (int q = 0; q < 4; ++q) {
    int64_t i = base + threadIdx.x * 4 + q;
    v[q] = i < n ? in[i] : 0;
    t += v[q];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  // Hillis-Steele inclusive scan over the 256 thread totals
  for (int o = 1; o < kBlock; o <<= 1) {
    int64_t a = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += a;
    __syncthreads();
  }
  int64_t run = bsum[blockIdx.x] + s[threadIdx.x] - t;
  for (int q = 0; q < 4; ++q) {
    int64_t i = base + threadIdx.x * 4 + q;
    if (i < n) out[i] = (int32_t)run;
    run += v[q];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = (int32_t)bsum[gridDim.x];
}

// Scatter entries of A into column buckets of A^T (order inside a bucket is
// fixed afterwards by k_sort_small / k_sort_long).
template <int L>
__global__ __launch_bounds__(kBlock) void k_tscatter(int64_t n, const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ col,
                                                     const int32_t* __restrict__ tptr, int32_t* fill,
                                                     int32_t* tcol, int32_t* tperm) {
  const int grp = threadIdx.x / L, lane = threadIdx.x % L;
  const int64_t stride = (int64_t)gridDim.x * (kBlock / L);
  for (int64_t row = (int64_t)blockIdx.x * (kBlock / L) + grp; row < n; row += stride) {
    for (int j = ptr[row] + lane; j < ptr[row + 1]; j += L) {
      const int c = col[j];
      const int d = tptr[c] + atomicAdd(fill + c, 1);
      tcol[d] = (int32_t)row;
      tperm[d] = j;
    }
  }
}

// Order inside an A^T row segment.  The key is the source entry index
// (tperm): entries of one column of A taken in CSR order come in row order,
// ties between repeated (row, column) entries by their position in the row,
// so the result is the stable column-major order of a counting-sort
// transpose, duplicates included.  One thread per segment of <= 16 entries
// (insertion sort); segments of 17..64 entries are listed for k_sort_mid,
// longer ones for k_sort_long.
// Appends v to list for the active lanes with pred set: one atomic per
// wavefront (a counter hit once per lane serialises at the L2).  Called by
// all active lanes together.
__device__ __forceinline__ void wave_append(bool pred, int32_t v, int32_t* list, int32_t* cnt) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  const int lane = threadIdx.x & 63, leader = __ffsll((unsigned long long)m) - 1;
  int32_t base = 0;
  if (lane == leader) base = atomicAdd(cnt, __popcll(m));
  base = __shfl(base, leader, 64);
  if (pred) list[base + __popcll(m & ((1ull << lane) - 1))] = v;
}

__global__ void k_sort_small(int64_t n, const int32_t* __restrict__ tptr, int32_t* tcol, int32_t* tperm,
                             int32_t* long_rows, int32_t* n_long, int32_t* mid_rows, int32_t* n_mid) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
    const int s = tptr[r], e = tptr[r + 1];
    wave_append(e - s > 64, (int32_t)r, long_rows, n_long);
    wave_append(e - s > 16 && e - s <= 64, (int32_t)r, mid_rows, n_mid);
    if (e - s > 16) continue;
    for (int a = s + 1; a < e; ++a) {
      const int32_t kc = tcol[a];
      const int32_t kp = tperm[a];
      int b = a - 1;
      while (b >= s && tperm[b] > kp) {
        tcol[b + 1] = tcol[b];
        tperm[b + 1] = tperm[b];
        --b;
      }
      tcol[b + 1] = kc;
      tperm[b + 1] = kp;
    }
  }
}

// One wavefront per listed segment of 17..64 entries: a 64-lane bitonic sort
// by tperm in registers (shuffles), entry l on lane l, the unused lanes
// holding the largest key.  Segments are taken four to a workgroup.
__global__ __launch_bounds__(kBlock) void k_sort_mid(int32_t nmid, const int32_t* __restrict__ tptr,
                                                     const int32_t* __restrict__ mid_rows, int32_t* tcol,
                                                     int32_t* tperm) {
  const int lane = threadIdx.x & 63;
  const int32_t q = (int32_t)blockIdx.x * (kBlock / 64) + (int32_t)(threadIdx.x >> 6);
  if (q >= nmid) return;  // the whole wavefront
  const int32_t r = mid_rows[q];
  const int s = tptr[r], len = tptr[r + 1] - s;
  int32_t key = lane < len ? tperm[s + lane] : 0x7fffffff;
  int32_t pay = lane < len ? tcol[s + lane] : 0;
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int32_t pk = __shfl_xor(key, j, 64);
      const int32_t pp = __shfl_xor(pay, j, 64);
      const bool lower = (lane & j) == 0, up = (lane & k) == 0;
      if (lower == up ? pk < key : pk > key) {
        key = pk;
        pay = pp;
      }
    }
  }
  if (lane < len) {
    tperm[s + lane] = key;
    tcol[s + lane] = pay;
  }
}

// Bitonic sort by tperm of [s, s + len), len <= kSortMax, in LDS by one
// workgroup of 1024 threads.
constexpr int kSortMax = 8192;
__device__ __forceinline__ void bitonic_segment(int64_t s, int len, int32_t* tcol, int32_t* tperm) {
  __shared__ int32_t kc[kSortMax];
  __shared__ int32_t kp[kSortMax];
  int P = 1;
  while (P < len) P <<= 1;
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    kp[i] = i < len ? tperm[s + i] : 0x7fffffff;
    kc[i] = i < len ? tcol[s + i] : 0;
  }
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          if ((kp[i] > kp[j]) == up) {
            int32_t a = kc[i]; kc[i] = kc[j]; kc[j] = a;
            int32_t b = kp[i]; kp[i] = kp[j]; kp[j] = b;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    tcol[s + i] = kc[i];
    tperm[s + i] = kp[i];
  }
}

// One workgroup per listed segment of 65..kSortMax entries; longer ones are
// left to k_sort_chunks + k_merge_runs.
__global__ __launch_bounds__(1024) void k_sort_long(const int32_t* __restrict__ tptr,
                                                    const int32_t* __restrict__ long_rows, int32_t* tcol,
                                                    int32_t* tperm) {
  const int32_t r = long_rows[blockIdx.x];
  const int s = tptr[r], len = tptr[r + 1] - s;
  if (len > kSortMax) return;
  bitonic_segment(s, len, tcol, tperm);
}

// One workgroup per chunk [cs[b], cs[b] + cl[b]) of a segment longer than
// kSortMax; chunks are aligned to kSortMax from the segment start.
__global__ __launch_bounds__(1024) void k_sort_chunks(const int32_t* __restrict__ cs, const int32_t* __restrict__ cl,
                                                      int32_t* tcol, int32_t* tperm) {
  bitonic_segment(cs[blockIdx.x], cl[blockIdx.x], tcol, tperm);
}

// One merge pass over the long segments: segment h holds len_h = off[h+1] -
// off[h] entries, at src + sb[h] (sb = the segment start in A^T, or off[h] in
// the flat staging buffer) and goes to dst + db[h].  Sorted runs of width w
// are merged pairwise: an entry's place is its index in its own run plus the
// number of keys in the partner run below its own (keys are distinct).
__global__ void k_merge_runs(int64_t total, int32_t nh, const int64_t* __restrict__ off,
                             const int64_t* __restrict__ sb, const int64_t* __restrict__ db, int64_t w,
                             const int32_t* __restrict__ sc, const int32_t* __restrict__ sp,
                             int32_t* __restrict__ dc, int32_t* __restrict__ dp) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    int lo = 0, hi = nh - 1;  // last h with off[h] <= e
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= e) lo = mid;
      else hi = mid - 1;
    }
    const int64_t i = e - off[lo], len = off[lo + 1] - off[lo];
    const int64_t run = i / w, pos = i - run * w, partner = run ^ 1;
    const int64_t pb = partner * w, plen = pb < len ? (len - pb < w ? len - pb : w) : 0;
    const int64_t s0 = sb[lo];
    const int32_t key = sp[s0 + i];
    int64_t a = 0, b = plen;  // count of partner keys < key
    while (a < b) {
      const int64_t mid = (a + b) >> 1;
      if (sp[s0 + pb + mid] < key) a = mid + 1;
      else b = mid;
    }
    const int64_t o = (run < partner ? run : partner) * w + pos + a;
    dp[db[lo] + o] = key;
    dc[db[lo] + o] = sc[s0 + i];
  }
}

template <class Tin, class Tout>
__global__ void k_gather(int64_t m, const int32_t* __restrict__ perm, const Tin* __restrict__ in,
                         Tout* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride)
    out[j] = (Tout)in[perm[j]];
}

// out[j] = col[j] - r for every entry j of row r (the int16 column offsets
// of a matrix whose offsets all lie in [-32767, 32767], checked by the host)
__global__ void k_col_offsets(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                              int16_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride)
    for (int32_t j = ptr[r], e = ptr[r + 1]; j < e; ++j) out[j] = (int16_t)(col[j] - r);
}

// Delta-coded columns (see k_spmv_stream): out[first] = col - r + kDeltaBias,
// out[j] = col[j] - col[j - 1] after it; *bad is set when a row is not
// sorted or a code does not fit 16 bits (the host then keeps int32 columns)
__global__ void k_col_delta(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                            uint16_t* __restrict__ out, int* __restrict__ bad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
    int64_t prev = r - kDeltaBias;
    for (int32_t j = ptr[r], e = ptr[r + 1]; j < e; ++j) {
      const int64_t d = (int64_t)col[j] - prev;
      if (d < 0 || d > 65535) *bad = 1;
      out[j] = (uint16_t)d;
      prev = col[j];
    }
  }
}

// out[0..m) = v: clears memory with a kernel (no memset node), for state
// that lives in memory the library does not allocate per call.
template <class T>
__global__ void k_fill(int64_t m, T* __restrict__ out, T v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) out[j] = v;
}

// Bandwidth references for ks_bench_stream: a plain grid-stride read of
// 16-byte chunks (one partial per wavefront, so the loads stay live) and a
// copy, non-temporal, with nothing else in the loop; the fill initialises
// the source.
using StreamChunk = Vec16<float>::type;
__global__ __launch_bounds__(kBlock) void k_stream_read(int64_t n16, const StreamChunk* __restrict__ x,
                                                        float* __restrict__ out) {
  float s = 0.f;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride) {
    const StreamChunk v = __builtin_nontemporal_load(x + i);
    s += (v[0] + v[1]) + (v[2] + v[3]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) out[blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)] = s;
}
__global__ __launch_bounds__(kBlock) void k_stream_copy(int64_t n16, const StreamChunk* __restrict__ x,
                                                        StreamChunk* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(x + i), y + i);
}
__global__ __launch_bounds__(kBlock) void k_stream_fill(int64_t n16, StreamChunk* __restrict__ y, float v) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const StreamChunk c = {v, v, v, v};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride) y[i] = c;
}

template <class Tin, class Tout>
__global__ void k_convert(int64_t m, const Tin* __restrict__ in, Tout* __restrict__ out, Gate g) {
  if (g.off()) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride)
    out[j] = (Tout)in[j];
}

}  // namespace ks
