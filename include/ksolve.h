/*
 * ksolve.h - C ABI of the MI355X (gfx950) Krylov solver library libksolve.so.
 *
 * This is the drop-in boundary for the Krylov hot path of ikuno/solver_cpp.
 * Every entry point below names the reference interface it replaces
 * (file:line in the reference tree).  Plain C types only: pointers, sizes,
 * fixed-size structs.  No exceptions cross the ABI, no call exits the
 * process; every call returns a ks_status.
 *
 * Pointer location: vector/matrix arguments are host pointers unless the
 * KS_PTR_DEVICE flag is passed, in which case they are device pointers on
 * the context's GPU (e.g. torch tensors' data_ptr()).
 *
 * Threading: a ks_ctx is single-threaded and stream-ordered (one HIP stream
 * per context).  Different contexts may be used from different threads.
 */
#ifndef KSOLVE_H_INCLUDED
#define KSOLVE_H_INCLUDED

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KS_ABI_VERSION 1

/* Status codes.  KS_OK / KS_NOT_CONVERGED keep the reference's solver exit
 * flags 0 / 2 (outerMethods.cpp:74-81, innerMethods.cpp:114-117); the rest
 * replace its exit(-1) paths (solver_collection.cpp:497-499,701-719,
 * cudaFunction.cu:903-925). */
typedef enum ks_status {
  KS_OK = 0,
  KS_NOT_CONVERGED = 2,
  KS_BREAKDOWN = 3,          /* non-finite or zero pivot in a recurrence */
  KS_ERR_BAD_INPUT = 10,     /* malformed CSR, unknown solver name, bad size */
  KS_ERR_HIP = 11,           /* HIP runtime error (message in ks_ctx_last_error) */
  KS_ERR_UNSUPPORTED = 12,   /* valid request this build does not implement */
  KS_ERR_NO_MEMORY = 13,
  KS_ERR_IO = 14             /* CRS text file missing / header mismatch */
} ks_status;

/* ks_params.ref_compat bits: reproduce reference behaviours that differ from
 * the textbook algorithm, so iteration counts and x match it bug-for-bug. */
#define KS_COMPAT_LAST_TERM  1u /* GMRES/VPGMRES exits inside a cycle assemble
                                   x = x0 + y[k-1]*V[k-1] only: Scalar_ax
                                   assigns (blas.cpp:211-224, gmres.cpp:291-298,
                                   311-318, vpgmres.cpp:798-805,818-825) */
#define KS_COMPAT_KSKIP_AS_WRITTEN 2u /* k-skip CG delta/eta/zeta updates as
                                   written at kskipcg.cpp:283-288 */
#define KS_COMPAT_VPCG_FR 4u    /* VPCG beta = (r',z')/(r,z) (vpcg.cpp:748-753);
                                   without it the flexible (Polak-Ribiere)
                                   beta = (z', r'-r)/(r,z) is used, which stays
                                   convergent under a variable inner solve */
#define KS_COMPAT_ALL 7u

/* flags for pointer arguments and options */
#define KS_PTR_DEVICE   1u  /* vector/matrix pointers are device pointers */
#define KS_X0_GIVEN     2u  /* ks_solve: use *x as the initial guess (the
                               reference always zeroes x: cg.cpp:79) */
#define KS_MAT_FP32_COPY 4u /* keep an FP32 copy of the values (mixed inner) */
#define KS_MAT_TRANSPOSE 8u /* build the explicit transposed CSR (BiCG, adjoint) */
#define KS_MAT_SPD      16u /* caller asserts A is symmetric positive definite */
#define KS_MAT_MERGE    32u /* merge-path SpMV for this matrix (with no
                               layout flag it is chosen automatically when the
                               longest row exceeds 256 entries and 16x the mean) */
#define KS_MAT_VECTOR   64u /* force the lanes-per-row vector SpMV */
/* CSR-stream tiles: by default the SpMV works on workgroup tiles (256 rows /
 * 2041 entries); KS_MAT_WAVE_TILES selects wavefront tiles (64 / 509) */
#define KS_MAT_WG_TILES 128u   /* workgroup tiles (the default) */
#define KS_MAT_WAVE_TILES 512u /* wavefront tiles */
#define KS_MAT_PANELS   1024u /* column panels: the SpMV runs as 10 passes over
                                 column slices of x, each small enough for an
                                 XCD's L2 (chosen automatically, with no layout
                                 flag, for n >= 2^21 rows when over half the
                                 entries lie more than n/8 columns from their
                                 row) */
#define KS_MAT_COL32    2048u /* keep int32 column indices in the SpMV.  Without
                                 it, a matrix whose entries all lie within
                                 32767 columns of their row (stencils, banded
                                 matrices) also gets int16 offsets col - row,
                                 which the workgroup-tile SpMV reads instead
                                 (2 bytes per entry less; same results).
                                 Otherwise, when A's (or A^T's) rows are
                                 sorted, each row's first column lies in
                                 [row - 49152, row + 16383], the gaps between
                                 its columns are below 65536 and no tile
                                 averages over 8 entries per row, that side
                                 gets delta-coded 16-bit columns instead
                                 (3-D stencils up to 221^3; same results) */
#define KS_MAT_COL_DELTA 4096u /* prefer the delta-coded 16-bit columns over
                                  int16 offsets where both apply (they read
                                  the same bytes; int16 offsets have no
                                  tile condition) */
#define KS_MAT_BORROW   256u /* with KS_PTR_DEVICE: use the caller's ptr/col/val
                                in place (no copy); they must outlive the
                                matrix, col and val must be 16-byte aligned,
                                and values changed in place are made
                                visible with ks_mat_update_values(m, val) */

/* precision selectors */
#define KS_F64 64
#define KS_F32 32

/* Solver parameters: the reference's flag surface (solver_collection.cpp:
 * 278-307, defaults :20-49).  ks_params_default() fills the defaults. */
typedef struct ks_params {
  char outer[16];            /* -S: cg cr gcr bicg gmres kskipcg kskipbicg
                                vpcg vpgcr vpgmres (case as reference) */
  char inner[16];            /* -s: "NO" or cg cr gcr bicg gmres kskipcg kskipbicg */
  int32_t outer_maxloop;     /* -L  10000 */
  int32_t outer_restart;     /* -R  1000 (GMRES: R-1 Arnoldi steps/cycle) */
  double outer_eps;          /* -E  1e-8 */
  int32_t outer_kskip;       /* -K  2 */
  int32_t outer_fix;         /* -F  2 */
  int32_t inner_maxloop;     /* -l  50 */
  int32_t inner_restart;     /* -r  1000 */
  double inner_eps;          /* -e  1e-1 */
  int32_t inner_kskip;       /* -k  2 */
  int32_t inner_fix;         /* -f  2 */
  int32_t inner_precision;   /* KS_F64 (reference) or KS_F32 (mixed) */
  uint32_t ref_compat;       /* KS_COMPAT_* bits */
  int32_t omp_threads;       /* -t  8: CPU-side reductions only */
  int32_t verbose;           /* -v */
  uint32_t flags;            /* KS_PTR_DEVICE | KS_X0_GIVEN */
  int32_t check_every;       /* CG-family: iterations enqueued per host check, 0=auto */
  int32_t profile;           /* 1: time every kernel class with hipEvents */
  int32_t reserved;
} ks_params;

/* kernel classes for ks_result.prof_* */
enum {
  KS_K_SPMV64 = 0, KS_K_SPMV32 = 1, KS_K_REDUCE = 2, KS_K_CGS = 3,
  KS_K_AXPY = 4, KS_K_SCALAR = 5, KS_K_BASIS = 6, KS_K_OTHER = 7, KS_K_NCLASS = 8
};

typedef struct ks_result {
  int32_t status;            /* ks_status of the solve */
  int32_t reserved;
  int64_t iters;             /* the `loop =` value the reference prints */
  int64_t steps;             /* outer Krylov steps performed */
  int64_t inner_steps;       /* inner Krylov steps performed (VP solvers) */
  int64_t spmv;              /* SpMV launches (all precisions) */
  int64_t n_hist;            /* history entries written */
  double relres;             /* residual estimate the stop test last saw */
  double relres_true;        /* ||b-Ax||/||b|| recomputed in FP64 */
  double log10_check;        /* log10(||b-Ax||/||b-Ax0||): Check_error (blas.cpp:266-301) */
  double seconds;            /* device-synchronised wall time of the solve */
  double bytes;              /* algorithmic HBM bytes of all launched kernels */
  double prof_ms[KS_K_NCLASS];      /* hipEvent time per kernel class (profile=1) */
  int64_t prof_launches[KS_K_NCLASS];
  double prof_bytes[KS_K_NCLASS];   /* algorithmic bytes per kernel class */
} ks_result;

typedef struct ks_ctx ks_ctx;
typedef struct ks_mat ks_mat;

/* ---- context ------------------------------------------------------------
 * Replaces cuda::cuda(times*, N[, k]) (cudaFunction.cu:270-346) and the
 * device part of collection::collection (solver_collection.cpp:16-100). */
int32_t ks_abi_version(void);
void ks_params_default(ks_params* p);
const char* ks_status_string(int32_t status);
ks_status ks_ctx_create(int32_t device, ks_ctx** out);
void ks_ctx_destroy(ks_ctx* ctx);
const char* ks_ctx_last_error(const ks_ctx* ctx);
/* hipStream_t the context launches on (as void*), for callers that order
 * their own work against it. */
void* ks_ctx_stream(ks_ctx* ctx);
ks_status ks_ctx_synchronize(ks_ctx* ctx);

/* ---- matrix -------------------------------------------------------------
 * Replaces collection::CRSAlloc_Part2 + CudaCopy_Part1 + transposeMatrix
 * (solver_collection.cpp:857-868,932-942,1021-1125) and cuda::CSR2CSC
 * (cudaFunction.cu:896-931): 0-based int32 CSR, N==M, values FP64.  The
 * library copies the arrays to HBM (owned), or with KS_MAT_BORROW uses the
 * caller's device arrays in place.  KS_MAT_TRANSPOSE builds A^T on
 * the device (count, scan, scatter, per-column sort), KS_MAT_FP32_COPY
 * keeps FP32 values for mixed-precision inner solves. */
ks_status ks_mat_create_csr(ks_ctx* ctx, int64_t n, int64_t nnz,
                            const int32_t* ptr, const int32_t* col,
                            const double* val, uint32_t flags, ks_mat** out);
/* Replace the values (same pattern); refreshes the FP32 copy and A^T values. */
ks_status ks_mat_update_values(ks_mat* m, const double* val, uint32_t flags);
/* Copy the device-built A^T (CSR of the transpose) out, for checking. */
ks_status ks_mat_get_transpose(ks_mat* m, int32_t* tptr, int32_t* tcol,
                               double* tval, uint32_t flags);
/* Destroy every matrix of a context before the context itself. */
void ks_mat_destroy(ks_mat* m);
/* n, nnz and the SpMV variant chosen: spmv_kind 0 CSR-stream tiles (workgroup
 * or wavefront), 1 lanes-per-row vector kernel, 2 merge path, 3 column panels;
 * lanes_per_row is the vector kernel's lanes per row (the smallest power of
 * two >= the mean row length, in [2, 64]), reported for every kind. */
ks_status ks_mat_info(const ks_mat* m, int64_t* n, int64_t* nnz,
                      int32_t* spmv_kind, int32_t* lanes_per_row);
/* Bytes per column index that the SpMV of A (trans = 0) or A^T (trans = 1)
 * reads, in a precision (KS_F64 / KS_F32): 2 with int16 offsets or
 * delta-coded columns (see KS_MAT_COL32), else 4.  No reference counterpart (the reference always
 * reads int32 columns, cudaFunction.cu:127-156); the bench uses it for the
 * SpMV's byte count. */
ks_status ks_mat_index_bytes(const ks_mat* m, int32_t trans, int32_t precision, int32_t* bytes);

/* ---- operators (the BLAS layer; tests and benchmarks call these) -------- */
/* y = A x or y = A^T x.  Replaces cuda::MtxVec_mult (cudaFunction.cu:613-638,
 * kernel at :127-156) and blas::MtxVec_mult (blas.cpp:43-129).  precision
 * KS_F64 (double x,y, FP64 values) or KS_F32 (float x,y, FP32 value copy). */
ks_status ks_spmv(ks_mat* m, int32_t trans, int32_t precision,
                  const void* x, void* y, uint32_t flags);
/* Time `reps` back-to-back SpMV launches with hipEvents on the context
 * stream; returns the mean launch duration in ms (x, y device pointers). */
ks_status ks_spmv_bench(ks_mat* m, int32_t trans, int32_t precision,
                        const void* x, void* y, int32_t reps, double* ms_per_launch);
/* Measurement only, no reference counterpart: the HBM rate (GB/s) that a
 * plain streaming kernel reaches on this device over `bytes` (a multiple of
 * 16, 16 KiB to 16 GiB, allocated and freed inside the call): kind 0 reads,
 * 1 copies (2 x bytes moved); the best of four grid sizes, each
 * timed over `reps` launches with hipEvents on the context stream.  The
 * bench reports it beside the SpMV's fraction of the published peak. */
ks_status ks_bench_stream(ks_ctx* ctx, int32_t kind, int64_t bytes, int32_t reps, double* gbps);
/* dot / norm: replaces blas::dot and blas::norm_2 (blas.cpp:31-41,149-209). */
ks_status ks_dot(ks_ctx* ctx, int64_t n, int32_t precision, const void* x,
                 const void* y, double* out, uint32_t flags);
ks_status ks_norm2(ks_ctx* ctx, int64_t n, int32_t precision, const void* x,
                   double* out, uint32_t flags);
/* out = a*x + y: replaces blas::Scalar_axy (blas.cpp:226-255). */
ks_status ks_axpy(ks_ctx* ctx, int64_t n, double a, const double* x,
                  const double* y, double* out, uint32_t flags);
/* Classical Gram-Schmidt step on a column-major n x (k+1) block V:
 * h[i] = (w, V_i) for i<=k, then w -= sum_i h[i] V_i, then
 * h[k+1] = ||w||.  Replaces the dot loop + blas::Gmres_sp_1
 * (gmres.cpp:354-377, blas.cpp:477-488). */
ks_status ks_cgs(ks_ctx* ctx, int64_t n, int32_t k, const double* V, double* w,
                 double* h, uint32_t flags);
/* Matrix-powers basis Ar[i]=A^{i+1} r (i<=2k), Ap[i]=A^{i+1} p (i<=2k+1),
 * column-major blocks: replaces cuda::Kskip_cg_bicg_base
 * (cudaFunction.cu:984-1026) / blas::Kskip_cg_bicg_base (blas.cpp:353-399). */
ks_status ks_kskip_basis(ks_mat* m, int32_t k, const double* r, const double* p,
                         double* Ar, double* Ap, uint32_t flags);
/* Block Gram products of the k-skip CG basis: delta[i]=(r,Ar_i) i<2k,
 * eta[i]=(r,Ap_i) i<2k+1, zeta[i]=(p,Ap_i) i<2k+2.  Replaces
 * blas::Kskip_cg_innerProduce (blas.cpp:401-435).  Outputs are host arrays. */
ks_status ks_kskip_gram(ks_ctx* ctx, int64_t n, int32_t k, const double* r,
                        const double* p, const double* Ar, const double* Ap,
                        double* delta, double* eta, double* zeta, uint32_t flags);
/* out[j] = alpha * u[row(j)] * v[col(j)] over the sparsity pattern of A
 * (SDDMM).  No reference counterpart: it gives the adjoint's d/d(values). */
ks_status ks_sddmm(ks_mat* m, double alpha, const double* u, const double* v,
                   double* out_vals, uint32_t flags);

/* ---- solvers ------------------------------------------------------------
 * Replaces outerMethods::outerSelect + <solver>::solve (outerMethods.cpp:
 * 20-82; cg.cpp:148-272, bicg.cpp:162-310, gmres.cpp:232-457, kskipcg.cpp:
 * 187-343, kskipbicg.cpp:207-385, gcr.cpp:178-347, cr.cpp:152-295,
 * vpcg.cpp:659-784, vpgmres.cpp:746-975) with the inner dispatch of
 * innerMethods::innerSelect (innerMethods.cpp:71-118).  x0 = 0 unless
 * KS_X0_GIVEN.  hist (host, may be NULL) receives (index, relres) pairs,
 * the lines the reference writes to ./output/<NAME>_his.txt; hist_cap is
 * the number of pairs it can hold. */
ks_status ks_solve(ks_ctx* ctx, ks_mat* m, const ks_params* p, const double* b,
                   double* x, double* hist, int64_t hist_cap, ks_result* res);

/* No-allocation variant of ks_solve.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory during a
 * call (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It runs plain FP64 CG (p->outer "cg", p->inner "NO")
 * with the same kernels and results as ks_solve, on device pointers only
 * (KS_PTR_DEVICE), with no residual history and no profiling, for matrices
 * of at most 1,048,576 rows (KS_ERR_BAD_INPUT beyond that).  Its vectors and
 * control state live in a fixed device array of about 24 MB that the library
 * holds on each GPU it runs on until the process ends; it calls no hipMalloc
 * or hipFree.  Calls on one GPU share that array: the library serialises
 * them within a process (one at a time per device), and each call keeps the
 * array's contents until the next. */
ks_status ks_solve_noalloc(ks_ctx* ctx, ks_mat* m, const ks_params* p, const double* b,
                           double* x, ks_result* res);

/* ---- adjoint ------------------------------------------------------------
 * No reference counterpart.  Given the forward solution x of A x = b and
 * g = dL/dx, solves A^T lambda = g with the solver named in p (cg for SPD,
 * bicg/gmres/vpgmres on the explicit transpose otherwise) and returns
 * grad_b = lambda and grad_vals[j] = -lambda[row(j)] * x[col(j)]. */
ks_status ks_adjoint(ks_ctx* ctx, ks_mat* m, const ks_params* p, const double* x,
                     const double* dLdx, double* grad_b, double* grad_vals,
                     ks_result* res);

/* ---- synthetic inputs (host, deterministic) -----------------------------
 * No reference counterpart (the reference only reads CRS text files).
 * Two-phase: pass NULL arrays to get n and nnz, then call again with arrays.
 * RNG: xoshiro256** seeded through splitmix64; normals by Box-Muller. */
/* dim 2: 5-point, dim 3: 7-point first-order upwind convection-diffusion on
 * an nx*ny*nz grid, Dirichlet boundary; c_d = 2*pe*vel_d/max|vel|;
 * A_ii = 2*dim + sum c_d, A_{i,i-e_d} = -1 - c_d, A_{i,i+e_d} = -1.  pe=0 and
 * diag=4 (2-D) / 6 (3-D) is the Poisson matrix. */
ks_status ks_gen_stencil(int32_t dim, int32_t nx, int32_t ny, int32_t nz,
                         double pe, const double* vel, int64_t* n, int64_t* nnz,
                         int32_t* ptr, int32_t* col, double* val);
/* 2-D graph Laplacian on nx*ny with per-edge conductivity exp(N(0,1)) drawn
 * from `seed` (Dirichlet boundary edges included).  edge_kappa (optional,
 * length (nx+1)*ny + nx*(ny+1)) receives the conductivities; edge_of_val
 * (optional, length nnz) maps each value slot to the edge it belongs to
 * (diagonal slots get -1). */
ks_status ks_gen_aniso2d(int32_t nx, int32_t ny, uint64_t seed, int64_t* n,
                         int64_t* nnz, int32_t* ptr, int32_t* col, double* val,
                         double* edge_kappa);
/* Assemble the aniso2d matrix values from given edge conductivities
 * (pattern from ks_gen_aniso2d). */
ks_status ks_aniso2d_values(int32_t nx, int32_t ny, const double* edge_kappa,
                            double* val);
/* Random nonsymmetric matrix: row lengths Zipf(s) on 1..cap (including the
 * diagonal), distinct uniform off-diagonal columns, values N(0,1), diagonal
 * 1.1*sum|a_ij| (1.0 for rows with no off-diagonal). */
ks_status ks_gen_zipf(int64_t n, double s, int32_t cap, uint64_t seed_len,
                      uint64_t seed_col, uint64_t seed_val, int64_t* nnz,
                      int32_t* ptr, int32_t* col, double* val);
/* dist 0: U(lo,hi), 1: N(0,1) */
ks_status ks_gen_vector(uint64_t seed, int32_t dist, double lo, double hi,
                        int64_t n, double* out);

/* ---- reference CRS text files (solver_collection.cpp:655-828) ----------- */
ks_status ks_crs_read_header(const char* dir, int64_t* n, int64_t* nnz);
ks_status ks_crs_read(const char* dir, int32_t* ptr, int32_t* col, double* val,
                      double* b, double* x);
ks_status ks_crs_write(const char* dir, int64_t n, int64_t nnz, const int32_t* ptr,
                       const int32_t* col, const double* val, const double* b,
                       const double* x);

/* ---- binary CSR sidecar of the CRS text files ----------------------------
 * No reference counterpart: a versioned binary copy of ColVal/Ptr/bx (the
 * Solver tool keeps it as <dir>/CRS.ksb with --bin-cache) that loads at
 * file-read speed instead of parsing text (collection::readMatrix,
 * solver_collection.cpp:726-828, fscanf per line).  The header records a
 * stamp of the text files it was made from (ks_crs_stamp: sizes and
 * modification times) and a checksum of the arrays; ks_crsbin_read rejects a
 * file whose size or checksum does not match (KS_ERR_IO). */
ks_status ks_crs_stamp(const char* dir, uint64_t* stamp);
ks_status ks_crsbin_write(const char* path, int64_t n, int64_t nnz, const int32_t* ptr,
                          const int32_t* col, const double* val, const double* b,
                          const double* x, uint64_t stamp);
ks_status ks_crsbin_read_header(const char* path, int64_t* n, int64_t* nnz,
                                uint64_t* stamp);
ks_status ks_crsbin_read(const char* path, int32_t* ptr, int32_t* col, double* val,
                         double* b, double* x);

#ifdef __cplusplus
}
#endif
#endif /* KSOLVE_H_INCLUDED */
