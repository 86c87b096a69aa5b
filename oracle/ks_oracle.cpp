// ks_oracle.cpp - CPU restatement of ikuno/solver_cpp's Krylov path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the GPU library is
// checked against and the CPU baseline bench.py times; nothing in the
// product (solver_cpp_amd/, libksolve.so) links, loads or calls it.
//
// It follows the reference's loop structure line by line (citations are
// file:line in the reference tree) in FP64, with two switches:
//   * ks_params.ref_compat (KS_COMPAT_*): reproduce the reference's
//     non-textbook behaviours, or use the textbook algorithm;
//   * ks_params.inner_precision 32: the VP inner solve runs on an FP32 copy
//     of A with FP32 vectors (SpMV accumulates in FP32, reductions and the
//     small dense GMRES work in FP64) - the arithmetic the GPU's mixed path
//     uses, so the two can be compared step for step.
//
// Reductions: the reference's dots are `omp parallel for schedule(static)
// reduction(+:tmp)` over T threads (blas.cpp:149-209, T = -t, default 8).
// Here the static partition is emulated exactly (libgomp's static schedule:
// the first n%T chunks get one extra element) and the T partial sums are
// combined in thread order, so results are deterministic for a given T.
// The reference's SpMV sums each row sequentially in CSR order
// (blas.cpp:56-62) without FMA: build with -ffp-contract=off.
//
// Parity pins of each function (closed forms, invariants, brute force) are
// in tests/test_oracle.py; the functions are listed with their pins in the
// header comment of that file.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/ksolve.h"

namespace {

using i64 = int64_t;

template <class T>
struct Csr {
  i64 n = 0, nnz = 0;
  const int32_t* ptr = nullptr;
  const int32_t* col = nullptr;
  const T* val = nullptr;
};

struct Cfg {
  int maxloop = 0;
  double eps = 0;
  int restart = 0;
  int kskip = 0;
  int fix = 0;
  bool inner = false;
  unsigned compat = 0;
  int nchunk = 8;
};

struct Out {
  int status = 2;
  i64 iters = 0, steps = 0, inner_steps = 0, spmv = 0;
  double relres = 0;
  std::vector<double> hist;  // (index, relres) pairs
  bool record = false;
  void his(double idx, double err) {
    if (record) {
      hist.push_back(idx);
      hist.push_back(err);
    }
  }
};

double g_kskip_oob = 0.0;  // value read past eta[] (kskipcg.cpp:284, j=2k-1)
// Trace of k-skip CG (for the tests' step-by-step checks): when enabled, each
// block appends its starting r and p (n values each), then (alpha, beta) of
// each of its k+1 inner steps.
bool g_kskip_trace_on = false;
std::vector<double> g_kskip_trace;

// ---------------------------------------------------------------- BLAS ----
// blas::dot (blas.cpp:149-162): emulated static partition over nchunk
// threads, partials combined in thread order.
template <class F>
double chunk_sum(i64 n, int nchunk, F term) {
  if (nchunk < 1) nchunk = 1;
  std::vector<double> part(nchunk, 0.0);
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nchunk; ++t) {
    i64 q = n / nchunk, r = n % nchunk;
    i64 lo = t * q + std::min<i64>(t, r);
    i64 hi = lo + q + (t < r ? 1 : 0);
    double s = 0.0;
    for (i64 i = lo; i < hi; ++i) s += term(i);
    part[t] = s;
  }
  double s = 0.0;
  for (int t = 0; t < nchunk; ++t) s += part[t];
  return s;
}

template <class T>
double dot(i64 n, const T* x, const T* y, int nc) {
  return chunk_sum(n, nc, [&](i64 i) { return (double)x[i] * (double)y[i]; });
}
// blas::norm_2 (blas.cpp:31-41)
template <class T>
double norm2(i64 n, const T* x, int nc) {
  return std::sqrt(chunk_sum(n, nc, [&](i64 i) { return (double)x[i] * (double)x[i]; }));
}

// blas::MtxVec_mult (blas.cpp:43-65): sequential row sums in CSR order,
// accumulated in the value type.
template <class T>
void spmv(const Csr<T>& A, const T* x, T* y) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < A.n; ++i) {
    T t = 0;
    for (int32_t j = A.ptr[i]; j < A.ptr[i + 1]; ++j) t += A.val[j] * x[A.col[j]];
    y[i] = t;
  }
}

// blas::Scalar_axy (blas.cpp:226-233): out = (a*x) + y, computed in double.
template <class T>
void axpy(i64 n, double a, const T* x, const T* y, T* out) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) out[i] = (T)((a * (double)x[i]) + (double)y[i]);
}
// blas::Vec_sub (blas.cpp:132-136)
template <class T>
void vsub(i64 n, const T* x, const T* y, T* out) {
  for (i64 i = 0; i < n; ++i) out[i] = x[i] - y[i];
}
// blas::Scalar_x_div_a (blas.cpp:257-264)
template <class T>
void xdiva(i64 n, const T* x, double a, T* out) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) out[i] = (T)((double)x[i] / a);
}

// blas::Hye (blas.cpp:303-314): back substitution on the rotated Hessenberg
// H (row stride ld), y zeroed on [0,size).
void hye(const double* h, int ld, double* y, const double* e, int size) {
  for (int i = 0; i < size; ++i) y[i] = 0.0;
  for (int i = size - 1; i >= 0; --i) {
    double t = 0.0;
    for (int j = i + 1; j < size; ++j) t += y[j] * h[(i)*ld + j];
    y[i] = (e[i] - t) / h[i * ld + i];
  }
}

// classical Gram-Schmidt (gmres.cpp:354-377, blas.cpp:477-488): all
// h_ik = (w, v_i) from the unmodified w, then w -= sum_i h_ik v_i with i
// sequential per element, then h_{k+1,k} = ||w||.
template <class T>
double cgs(i64 n, int k, const T* V, T* w, double* hcol, int ld, int nc) {
  for (int i = 0; i <= k; ++i) hcol[i * ld] = dot(n, w, V + (i64)i * n, nc);
#pragma omp parallel for schedule(static)
  for (i64 j = 0; j < n; ++j) {
    double t = (double)w[j];
    for (int i = 0; i <= k; ++i) t -= hcol[i * ld] * (double)V[(i64)i * n + j];
    w[j] = (T)t;
  }
  return norm2(n, w, nc);
}

// --------------------------------------------------------------- solvers --
template <class T>
struct Ws {
  const Csr<T>* A = nullptr;
  const Csr<T>* AT = nullptr;  // explicit transpose (BiCG family)
};

// cg::solve (cg.cpp:148-240)
template <class T>
void cg(const Ws<T>& W, const T* b, T* x, const Cfg& c, Out& o) {
  const Csr<T>& A = *W.A;
  i64 n = A.n;
  std::vector<T> r(n), p(n), mv(n);
  double bnorm = norm2(n, b, c.nchunk);
  spmv(A, x, mv.data()); o.spmv++;
  vsub(n, b, mv.data(), r.data());
  p = r;
  double rr = dot(n, r.data(), r.data(), c.nchunk);
  long loop;
  double err = 0;
  for (loop = 1; loop <= c.maxloop; ++loop) {
    err = norm2(n, r.data(), c.nchunk) / bnorm;
    o.his(loop, err);
    if (err <= c.eps) { o.status = 0; break; }
    spmv(A, p.data(), mv.data()); o.spmv++;
    double d = dot(n, p.data(), mv.data(), c.nchunk);
    double alpha = rr / d;
    axpy(n, alpha, p.data(), x, x);
    axpy(n, -alpha, mv.data(), r.data(), r.data());
    double rr2 = dot(n, r.data(), r.data(), c.nchunk);
    double beta = rr2 / rr;
    rr = rr2;
    axpy(n, beta, p.data(), r.data(), p.data());
    o.steps++;
  }
  o.iters = loop;
  o.relres = err;
}

// cr::solve (cr.cpp:152-280)
template <class T>
void cr(const Ws<T>& W, const T* b, T* x, const Cfg& c, Out& o) {
  const Csr<T>& A = *W.A;
  i64 n = A.n;
  std::vector<T> r(n), p(n), q(n), s(n);
  double bnorm = norm2(n, b, c.nchunk);
  spmv(A, x, q.data()); o.spmv++;
  vsub(n, b, q.data(), r.data());
  p = r;
  spmv(A, p.data(), q.data()); o.spmv++;
  s = q;
  double rs = dot(n, r.data(), s.data(), c.nchunk);
  long loop;
  double err = 0;
  for (loop = 1; loop <= c.maxloop; ++loop) {
    err = norm2(n, r.data(), c.nchunk) / bnorm;
    o.his(loop, err);
    if (err <= c.eps) { o.status = 0; break; }
    double qq = dot(n, q.data(), q.data(), c.nchunk);
    double alpha = rs / qq;
    axpy(n, alpha, p.data(), x, x);
    axpy(n, -alpha, q.data(), r.data(), r.data());
    spmv(A, r.data(), s.data()); o.spmv++;
    double rs2 = dot(n, r.data(), s.data(), c.nchunk);
    double beta = rs2 / rs;
    rs = rs2;
    axpy(n, beta, p.data(), r.data(), p.data());
    axpy(n, beta, q.data(), s.data(), q.data());
    o.steps++;
  }
  o.iters = loop;
  o.relres = err;
}

// bicg::solve (bicg.cpp:162-276): shadow r* = r0, A^T p* on the explicit
// transpose (bicg.cpp:245-254).
template <class T>
void bicg(const Ws<T>& W, const T* b, T* x, const Cfg& c, Out& o) {
  const Csr<T>& A = *W.A;
  i64 n = A.n;
  std::vector<T> r(n), rs(n), p(n), ps(n), mv(n);
  double bnorm = norm2(n, b, c.nchunk);
  spmv(A, x, mv.data()); o.spmv++;
  vsub(n, b, mv.data(), r.data());
  rs = r; p = r; ps = rs;
  double rr = dot(n, rs.data(), r.data(), c.nchunk);
  long loop;
  double err = 0;
  for (loop = 1; loop <= c.maxloop; ++loop) {
    err = norm2(n, r.data(), c.nchunk) / bnorm;
    o.his(loop, err);
    if (err <= c.eps) { o.status = 0; break; }
    spmv(A, p.data(), mv.data()); o.spmv++;
    double d = dot(n, ps.data(), mv.data(), c.nchunk);
    double alpha = rr / d;
    axpy(n, alpha, p.data(), x, x);
    axpy(n, -alpha, mv.data(), r.data(), r.data());
    spmv(*W.AT, ps.data(), mv.data()); o.spmv++;
    axpy(n, -alpha, mv.data(), rs.data(), rs.data());
    double rr2 = dot(n, rs.data(), r.data(), c.nchunk);
    double beta = rr2 / rr;
    rr = rr2;
    axpy(n, beta, p.data(), r.data(), p.data());
    axpy(n, beta, ps.data(), rs.data(), ps.data());
    o.steps++;
  }
  o.iters = loop;
  o.relres = err;
}

// Preconditioner hook for flexible GMRES / VPCG: z = M(v).
struct Precond {
  virtual void apply(const double* v, double* z, Out& o) = 0;
  virtual ~Precond() {}
};

// gmres::solve (gmres.cpp:232-423); with M != null this is vpgmres::solve
// (vpgmres.cpp:746-944): z_k = M(v_k) (:840), w = A z_k (:848), Z[k] = z_k
// (:861) and x is assembled from Z (:934-938).
//   * R = restart gives R-1 Arnoldi steps per cycle (:277);
//   * the iteration limit trips at count+1 >= maxloop before the eps test;
//   * converged / limit exits assemble x0 + sum_{i<k} y_i V_i, or with
//     KS_COMPAT_LAST_TERM only y_{k-1} V_{k-1} (Scalar_ax assigns, :291-298);
//   * the restart assembly sums i<R with yvec[R-1] never written by Hye
//     (gmres.cpp:409-417, undefined in the reference); taken as 0 here, which
//     is what vpgmres's zeroed yvec gives (vpgmres.cpp:408).
template <class T>
void gmres(const Ws<T>& W, const T* b, T* x, const Cfg& c, Out& o, Precond* M) {
  const Csr<T>& A = *W.A;
  i64 n = A.n;
  const int R = c.restart;
  if (R < 2) { o.status = KS_ERR_BAD_INPUT; return; }
  const int ld = R;
  std::vector<T> r(n), w(n), tmp(n), V((size_t)n * (R + 1));
  std::vector<T> Z;
  std::vector<double> zd, vd;
  if (M) { Z.assign((size_t)n * (R + 1), T(0)); zd.resize(n); vd.resize(n); }
  std::vector<double> H((size_t)(R + 1) * ld, 0.0), e(R, 0.0), y(R, 0.0), cs(R, 0.0), sn(R, 0.0);
  std::vector<T> x0(n, T(0));
  for (i64 j = 0; j < n; ++j) x0[j] = x[j];
  double bnorm = norm2(n, b, c.nchunk);
  bool exit_ok = false, over = false;
  long count = 0;
  double err = 0;
  const unsigned last_term = c.compat & KS_COMPAT_LAST_TERM;
  auto assemble = [&](int k) {
    const T* B = M ? Z.data() : V.data();
    hye(H.data(), ld, y.data(), e.data(), k);
    for (i64 j = 0; j < n; ++j) tmp[j] = 0;
    if (last_term) {
      if (k > 0)
        for (i64 j = 0; j < n; ++j) tmp[j] = (T)(y[k - 1] * (double)B[(i64)(k - 1) * n + j]);
    } else {
      for (int i = 0; i < k; ++i)
        for (i64 j = 0; j < n; ++j) tmp[j] = (T)((double)tmp[j] + y[i] * (double)B[(i64)i * n + j]);
    }
    for (i64 j = 0; j < n; ++j) x[j] = x0[j] + tmp[j];
  };
  for (count = 0; count < c.maxloop;) {
    spmv(A, x, w.data()); o.spmv++;
    vsub(n, b, w.data(), r.data());
    double t = norm2(n, r.data(), c.nchunk);
    xdiva(n, r.data(), t, V.data());
    std::fill(e.begin(), e.end(), 0.0);
    e[0] = t;
    for (int k = 0; k < R - 1; ++k) {
      err = std::fabs(e[k]) / bnorm;
      o.his(count + 1, err);
      if (count + 1 >= c.maxloop) { assemble(k); over = true; break; }
      if (err <= c.eps) { assemble(k); exit_ok = true; break; }
      const T* vk = V.data() + (i64)k * n;
      if (M) {
        for (i64 j = 0; j < n; ++j) vd[j] = (double)vk[j];
        M->apply(vd.data(), zd.data(), o);
        T* zk = Z.data() + (i64)k * n;
        for (i64 j = 0; j < n; ++j) zk[j] = (T)zd[j];
        spmv(A, zk, w.data()); o.spmv++;
      } else {
        spmv(A, vk, w.data()); o.spmv++;
      }
      double hn = cgs(n, k, V.data(), w.data(), H.data() + k, ld, c.nchunk);
      H[(k + 1) * ld + k] = hn;
      xdiva(n, w.data(), hn, V.data() + (i64)(k + 1) * n);
      for (int i = 0; i <= k - 1; ++i) {
        double t1 = H[i * ld + k], t2 = H[(i + 1) * ld + k];
        H[i * ld + k] = cs[i] * t1 - sn[i] * t2;
        H[(i + 1) * ld + k] = sn[i] * t1 + cs[i] * t2;
      }
      double hk = H[k * ld + k], hk1 = H[(k + 1) * ld + k];
      double alpha = std::sqrt(hk * hk + hk1 * hk1);
      cs[k] = hk / alpha;
      sn[k] = -hk1 / alpha;
      e[k + 1] = sn[k] * e[k];
      e[k] = cs[k] * e[k];
      H[k * ld + k] = cs[k] * H[k * ld + k] - sn[k] * H[(k + 1) * ld + k];
      H[(k + 1) * ld + k] = 0.0;
      count++;
      o.steps++;
    }
    if (exit_ok || over) break;
    // restart assembly over i < R (y[R-1] = 0, see above)
    hye(H.data(), ld, y.data(), e.data(), R - 1);
    y[R - 1] = 0.0;
    const T* B = M ? Z.data() : V.data();
    for (i64 j = 0; j < n; ++j) tmp[j] = 0;
    for (int i = 0; i < R; ++i)
      for (i64 j = 0; j < n; ++j) tmp[j] = (T)((double)tmp[j] + y[i] * (double)B[(i64)i * n + j]);
    for (i64 j = 0; j < n; ++j) x[j] = x0[j] + tmp[j];
    for (i64 j = 0; j < n; ++j) x0[j] = x[j];
  }
  o.status = exit_ok ? 0 : 2;
  o.iters = count + 1;
  o.relres = err;
}

// gcr::solve (gcr.cpp:178-313) with blas::Gcr_sp_1 (blas.cpp:526-543) in its
// race-free form: beta_i = -(Av, q_i)/qq_i for every i <= k, then
// p_{k+1} = sum_i beta_i p_i, q_{k+1} = sum_i beta_i q_i (i sequential),
// then p_{k+1} += r, q_{k+1} += Av.  loop starts at 1 (gcr.cpp:41).
template <class T>
void gcr(const Ws<T>& W, const T* b, T* x, const Cfg& c, Out& o) {
  const Csr<T>& A = *W.A;
  i64 n = A.n;
  const int R = c.restart;
  if (R < 1) { o.status = KS_ERR_BAD_INPUT; return; }
  std::vector<T> r(n), Av(n), P((size_t)n * R), Q((size_t)n * R);
  std::vector<double> qq(R), beta(R);
  double bnorm = norm2(n, b, c.nchunk);
  long loop = 1;
  bool out = false;
  double err = 0;
  o.status = 2;
  while (loop < c.maxloop) {
    spmv(A, x, Av.data()); o.spmv++;
    vsub(n, b, Av.data(), r.data());
    for (i64 j = 0; j < n; ++j) P[j] = r[j];
    spmv(A, P.data(), Q.data()); o.spmv++;
    for (int k = 0; k < R; ++k) {
      err = norm2(n, r.data(), c.nchunk) / bnorm;
      o.his(loop, err);
      if (err <= c.eps) { o.status = 0; out = true; break; }
      if (loop >= c.maxloop) { o.status = 2; out = true; break; }
      loop++;
      T* qk = Q.data() + (i64)k * n;
      T* pk = P.data() + (i64)k * n;
      qq[k] = dot(n, qk, qk, c.nchunk);
      double alpha = dot(n, r.data(), qk, c.nchunk) / qq[k];
      axpy(n, alpha, pk, x, x);
      o.steps++;
      if (k == R - 1) break;
      axpy(n, -alpha, qk, r.data(), r.data());
      spmv(A, r.data(), Av.data()); o.spmv++;
      T* pn = P.data() + (i64)(k + 1) * n;
      T* qn = Q.data() + (i64)(k + 1) * n;
      for (i64 j = 0; j < n; ++j) { pn[j] = 0; qn[j] = 0; }
      for (int i = 0; i <= k; ++i) beta[i] = -dot(n, Av.data(), Q.data() + (i64)i * n, c.nchunk) / qq[i];
      for (int i = 0; i <= k; ++i) {
        axpy(n, beta[i], P.data() + (i64)i * n, pn, pn);
        axpy(n, beta[i], Q.data() + (i64)i * n, qn, qn);
      }
      for (i64 j = 0; j < n; ++j) { pn[j] = r[j] + pn[j]; qn[j] = Av[j] + qn[j]; }
    }
    if (out) break;
  }
  o.iters = loop;
  o.relres = err;
}

// blas::Kskip_cg_bicg_base (blas.cpp:353-399): Ar_i = A^{i+1} r (i <= 2k),
// Ap_i = A^{i+1} p (i <= 2k+1); r and p products share one row pass.
template <class T>
void kskip_basis(const Csr<T>& A, int k, const T* r, const T* p, T* Ar, T* Ap) {
  i64 n = A.n;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) {
    T t1 = 0, t2 = 0;
    for (int32_t j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
      t1 += A.val[j] * r[A.col[j]];
      t2 += A.val[j] * p[A.col[j]];
    }
    Ar[i] = t1;
    Ap[i] = t2;
  }
  for (int ii = 1; ii < 2 * k + 2; ++ii) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {
      T t1 = 0, t2 = 0;
      for (int32_t j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
        if (ii < 2 * k + 1) t1 += A.val[j] * Ar[(i64)(ii - 1) * n + A.col[j]];
        t2 += A.val[j] * Ap[(i64)(ii - 1) * n + A.col[j]];
      }
      if (ii < 2 * k + 1) Ar[(i64)ii * n + i] = t1;
      Ap[(i64)ii * n + i] = t2;
    }
  }
}

// blas::Kskip_cg_innerProduce (blas.cpp:401-435): each column sum is
// sequential over j (the OpenMP loop is over columns).
template <class T>
void kskip_cg_gram(i64 n, int k, const T* r, const T* p, const T* Ar, const T* Ap,
                   double* delta, double* eta, double* zeta) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < 2 * k + 2; ++i) {
    double t1 = 0, t2 = 0, t3 = 0;
    for (i64 j = 0; j < n; ++j) {
      if (i < 2 * k) t1 += (double)r[j] * (double)Ar[(i64)i * n + j];
      if (i < 2 * k + 1) t2 += (double)r[j] * (double)Ap[(i64)i * n + j];
      t3 += (double)p[j] * (double)Ap[(i64)i * n + j];
    }
    if (i < 2 * k) delta[i] = t1;
    if (i < 2 * k + 1) eta[i] = t2;
    zeta[i] = t3;
  }
}

// blas::Kskip_bicg_innerProduce (blas.cpp:437-475)
template <class T>
void kskip_bicg_gram(i64 n, int k, const T* rs, const T* ps, const T* Ar, const T* Ap,
                     double* theta, double* eta, double* rho, double* phi) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < 2 * k + 2; ++i) {
    double t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    for (i64 j = 0; j < n; ++j) {
      if (i < 2 * k) t1 += (double)rs[j] * (double)Ar[(i64)i * n + j];
      if (i < 2 * k + 1) {
        t2 += (double)rs[j] * (double)Ap[(i64)i * n + j];
        t3 += (double)ps[j] * (double)Ar[(i64)i * n + j];
      }
      t4 += (double)ps[j] * (double)Ap[(i64)i * n + j];
    }
    if (i < 2 * k) theta[i] = t1;
    if (i < 2 * k + 1) { eta[i] = t2; rho[i] = t3; }
    phi[i] = t4;
  }
}

// kskipcg::solve (kskipcg.cpp:187-311).  Scalar recurrences: as written
// (:283-288) with KS_COMPAT_KSKIP_AS_WRITTEN, else the derivation from
// delta_j=(r,A^{j+1}r), eta_j=(r,A^{j+1}p), zeta_j=(p,A^{j+1}p):
//   delta' = delta_j - 2a eta_{j+1} + a^2 zeta_{j+2}
//   eta'   = delta'  + b eta_j - a b zeta_{j+1}
//   zeta'  = eta'    + b eta_j - a b zeta_{j+1} + b^2 zeta_j
// As written, j = 2k-1 reads eta[2k+1], one past the array (:284); that
// value never reaches alpha or beta (tests check invariance to it).
template <class T>
void kskipcg(const Ws<T>& W, const T* b, T* x, const Cfg& c, Out& o) {
  const Csr<T>& A = *W.A;
  i64 n = A.n;
  const int k = c.kskip;
  if (k < 0) { o.status = KS_ERR_BAD_INPUT; return; }
  std::vector<T> r(n), p(n), Av(n), Ar((size_t)(2 * k + 1) * n), Ap((size_t)(2 * k + 2) * n);
  std::vector<double> delta(2 * k + 1), eta(2 * k + 2), zeta(2 * k + 2);
  spmv(A, x, Av.data()); o.spmv++;
  vsub(n, b, Av.data(), r.data());
  p = r;
  double bnorm = norm2(n, b, c.nchunk);
  const bool as_written = c.compat & KS_COMPAT_KSKIP_AS_WRITTEN;
  long nloop;
  double err = 0;
  for (nloop = 0; nloop < c.maxloop; nloop += (k + 1)) {
    err = norm2(n, r.data(), c.nchunk) / bnorm;
    o.his(nloop + 1, err);
    if (err <= c.eps) { o.status = 0; break; }
    kskip_basis(A, k, r.data(), p.data(), Ar.data(), Ap.data());
    o.spmv += 4 * k + 3;
    double gamma = dot(n, r.data(), r.data(), c.nchunk);
    kskip_cg_gram(n, k, r.data(), p.data(), Ar.data(), Ap.data(), delta.data(), eta.data(), zeta.data());
    eta[2 * k + 1] = g_kskip_oob;
    if (g_kskip_trace_on) {
      for (i64 j = 0; j < n; ++j) g_kskip_trace.push_back((double)r[j]);
      for (i64 j = 0; j < n; ++j) g_kskip_trace.push_back((double)p[j]);
    }
    for (long il = nloop; il <= nloop + k; ++il) {
      double alpha = gamma / zeta[0];
      double beta = alpha * zeta[1] / zeta[0] - 1.0;
      if (g_kskip_trace_on) {
        g_kskip_trace.push_back(alpha);
        g_kskip_trace.push_back(beta);
      }
      if (c.fix == 1) {
        gamma = beta * gamma;
      } else if (c.fix == 2) {
        double t0 = gamma - alpha * eta[0];
        double t1 = eta[0] - alpha * zeta[1];
        gamma = t0 - alpha * t1;
      }
      int jm = 2 * k - 2 * (int)(il - nloop);
      for (int j = 0; j < jm; ++j) {
        if (as_written) {
          delta[j] = delta[j] - 2 * alpha * eta[j + 1] + alpha * alpha * eta[j + 2];
          double eta_old = eta[j];
          eta[j] = delta[j] + beta * zeta[j + 1] - alpha * beta * zeta[j + 1];
          zeta[j] = eta[j + 1] + beta * eta_old + beta * beta * zeta[j] - alpha * beta * zeta[j + 1];
        } else {
          double dn = delta[j] - 2 * alpha * eta[j + 1] + alpha * alpha * zeta[j + 2];
          double en = dn + beta * eta[j] - alpha * beta * zeta[j + 1];
          double zn = en + beta * eta[j] - alpha * beta * zeta[j + 1] + beta * beta * zeta[j];
          delta[j] = dn;
          eta[j] = en;
          zeta[j] = zn;
        }
      }
      spmv(A, p.data(), Av.data()); o.spmv++;
      axpy(n, alpha, p.data(), x, x);
      axpy(n, -alpha, Av.data(), r.data(), r.data());
      axpy(n, beta, p.data(), r.data(), p.data());
      o.steps++;
    }
  }
  o.iters = nloop - k + 1;
  o.relres = err;
}

// kskipBicg::solve (kskipbicg.cpp:207-352)
template <class T>
void kskipbicg(const Ws<T>& W, const T* b, T* x, const Cfg& c, Out& o) {
  const Csr<T>& A = *W.A;
  i64 n = A.n;
  const int k = c.kskip;
  if (k < 0) { o.status = KS_ERR_BAD_INPUT; return; }
  std::vector<T> r(n), p(n), rs(n), ps(n), Av(n), Ar((size_t)(2 * k + 1) * n), Ap((size_t)(2 * k + 2) * n);
  std::vector<double> theta(2 * k + 1), eta(2 * k + 2), rho(2 * k + 2), phi(2 * k + 2);
  spmv(A, x, Av.data()); o.spmv++;
  vsub(n, b, Av.data(), r.data());
  p = r; rs = r; ps = rs;
  double bnorm = norm2(n, b, c.nchunk);
  long nloop;
  double err = 0;
  for (nloop = 0; nloop < c.maxloop; nloop += (k + 1)) {
    err = norm2(n, r.data(), c.nchunk) / bnorm;
    o.his(nloop + 1, err);
    if (err <= c.eps) { o.status = 0; break; }
    kskip_basis(A, k, r.data(), p.data(), Ar.data(), Ap.data());
    o.spmv += 4 * k + 3;
    double gamma = dot(n, rs.data(), r.data(), c.nchunk);
    kskip_bicg_gram(n, k, rs.data(), ps.data(), Ar.data(), Ap.data(), theta.data(), eta.data(), rho.data(), phi.data());
    for (long il = nloop; il <= nloop + k; ++il) {
      double alpha = gamma / phi[0];
      double beta = 1.0 - (eta[0] + rho[0] - alpha * phi[1]) / phi[0];
      gamma = gamma * beta;
      int jm = 2 * k - 2 * (int)(il - nloop);
      for (int j = 0; j < jm; ++j) {
        theta[j] = theta[j] - alpha * eta[j + 1] - alpha * rho[j + 1] + alpha * alpha * phi[j + 2];
        double t = theta[j] - alpha * beta * phi[j + 1];
        eta[j] = t + beta * eta[j];
        rho[j] = t + beta * rho[j];
        phi[j] = eta[j] + rho[j] - theta[j] + beta * beta * phi[j];
      }
      spmv(A, p.data(), Av.data()); o.spmv++;
      axpy(n, alpha, p.data(), x, x);
      axpy(n, -alpha, Av.data(), r.data(), r.data());
      spmv(*W.AT, ps.data(), Av.data()); o.spmv++;
      axpy(n, -alpha, Av.data(), rs.data(), rs.data());
      axpy(n, beta, p.data(), r.data(), p.data());
      axpy(n, beta, ps.data(), rs.data(), ps.data());
      o.steps++;
    }
  }
  o.iters = nloop + 1;
  o.relres = err;
}

// vpcg::solve (vpcg.cpp:659-760): z = M(r) from z = 0 each time.
void vpcg(const Ws<double>& W, const double* b, double* x, const Cfg& c, Out& o, Precond* M) {
  const Csr<double>& A = *W.A;
  i64 n = A.n;
  std::vector<double> r(n), p(n), z(n), mv(n);
  double bnorm = norm2(n, b, c.nchunk);
  spmv(A, x, mv.data()); o.spmv++;
  vsub(n, b, mv.data(), r.data());
  M->apply(r.data(), z.data(), o);
  p = z;
  double rr = dot(n, r.data(), z.data(), c.nchunk);
  long loop;
  double err = 0;
  for (loop = 1; loop <= c.maxloop; ++loop) {
    err = norm2(n, r.data(), c.nchunk) / bnorm;
    o.his(loop, err);
    if (err <= c.eps) { o.status = 0; break; }
    spmv(A, p.data(), mv.data()); o.spmv++;
    double d = dot(n, p.data(), mv.data(), c.nchunk);
    double alpha = rr / d;
    axpy(n, alpha, p.data(), x, x);
    axpy(n, -alpha, mv.data(), r.data(), r.data());
    std::fill(z.begin(), z.end(), 0.0);
    M->apply(r.data(), z.data(), o);
    double rr2 = dot(n, r.data(), z.data(), c.nchunk);
    double beta;
    if (c.compat & KS_COMPAT_VPCG_FR) {
      beta = rr2 / rr;  // vpcg.cpp:753
    } else {
      // flexible: (z', r' - r) = -alpha (z', A p) since r' - r = -alpha A p
      beta = -alpha * dot(n, z.data(), mv.data(), c.nchunk) / rr;
    }
    rr = rr2;
    axpy(n, beta, p.data(), z.data(), p.data());
    o.steps++;
  }
  o.iters = loop;
  o.relres = err;
}

// vpgcr::solve (vpgcr.cpp:693-860): GCR with p_k = M(r) in place of r.
void vpgcr(const Ws<double>& W, const double* b, double* x, const Cfg& c, Out& o, Precond* M) {
  const Csr<double>& A = *W.A;
  i64 n = A.n;
  const int R = c.restart;
  if (R < 1) { o.status = KS_ERR_BAD_INPUT; return; }
  std::vector<double> r(n), Av(n), z(n), P((size_t)n * R), Q((size_t)n * R), qq(R), beta(R);
  double bnorm = norm2(n, b, c.nchunk);
  long loop = 1;
  bool out = false;
  double err = 0;
  o.status = 2;
  while (loop < c.maxloop) {
    spmv(A, x, Av.data()); o.spmv++;
    vsub(n, b, Av.data(), r.data());
    std::fill(z.begin(), z.end(), 0.0);
    M->apply(r.data(), z.data(), o);
    for (i64 j = 0; j < n; ++j) P[j] = z[j];
    spmv(A, P.data(), Q.data()); o.spmv++;
    for (int k = 0; k < R; ++k) {
      err = norm2(n, r.data(), c.nchunk) / bnorm;
      o.his(loop, err);
      if (err <= c.eps) { o.status = 0; out = true; break; }
      if (loop >= c.maxloop) { o.status = 2; out = true; break; }
      loop++;
      double* qk = Q.data() + (i64)k * n;
      double* pk = P.data() + (i64)k * n;
      qq[k] = dot(n, qk, qk, c.nchunk);
      double alpha = dot(n, r.data(), qk, c.nchunk) / qq[k];
      axpy(n, alpha, pk, x, x);
      o.steps++;
      if (k == R - 1) break;
      axpy(n, -alpha, qk, r.data(), r.data());
      std::fill(z.begin(), z.end(), 0.0);
      M->apply(r.data(), z.data(), o);
      spmv(A, z.data(), Av.data()); o.spmv++;
      double* pn = P.data() + (i64)(k + 1) * n;
      double* qn = Q.data() + (i64)(k + 1) * n;
      for (i64 j = 0; j < n; ++j) { pn[j] = 0; qn[j] = 0; }
      for (int i = 0; i <= k; ++i) beta[i] = -dot(n, Av.data(), Q.data() + (i64)i * n, c.nchunk) / qq[i];
      for (int i = 0; i <= k; ++i) {
        axpy(n, beta[i], P.data() + (i64)i * n, pn, pn);
        axpy(n, beta[i], Q.data() + (i64)i * n, qn, qn);
      }
      for (i64 j = 0; j < n; ++j) { pn[j] = z[j] + pn[j]; qn[j] = Av[j] + qn[j]; }
    }
    if (out) break;
  }
  o.iters = loop;
  o.relres = err;
}

// ---------------------------------------------------------- dispatch -----
enum Name { NONE, CG, BICG, CR, GCR, GMRES, KSKIPCG, KSKIPCR, KSKIPBICG, VPCG, VPBICG, VPCR, VPGCR, VPGMRES };

// collection::string2enum (solver_collection.cpp:238-267): lower or upper case.
Name name_of(const char* s) {
  std::string t(s);
  static const char* lo[] = {"", "cg", "bicg", "cr", "gcr", "gmres", "kskipcg", "kskipcr", "kskipbicg",
                             "vpcg", "vpbicg", "vpcr", "vpgcr", "vpgmres"};
  static const char* up[] = {"", "CG", "BICG", "CR", "GCR", "GMRES", "KSKIPCG", "KSKIPCR", "KSKIPBICG",
                             "VPCG", "VPBICG", "VPCR", "VPGCR", "VPGMRES"};
  for (int i = 1; i <= 13; ++i)
    if (t == lo[i] || t == up[i]) return (Name)i;
  return NONE;
}

void transpose(i64 n, i64 nnz, const int32_t* ptr, const int32_t* col, const double* val,
               int32_t* tptr, int32_t* tcol, double* tval) {
  // stable counting sort by column: A^T rows = A columns, row order kept
  std::vector<i64> cnt(n + 1, 0);
  for (i64 j = 0; j < nnz; ++j) cnt[col[j] + 1]++;
  for (i64 i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  for (i64 i = 0; i <= n; ++i) tptr[i] = (int32_t)cnt[i];
  std::vector<i64> pos(cnt.begin(), cnt.end() - 1);
  for (i64 i = 0; i < n; ++i)
    for (int32_t j = ptr[i]; j < ptr[i + 1]; ++j) {
      i64 d = pos[col[j]]++;
      tcol[d] = (int32_t)i;
      tval[d] = val[j];
    }
}

template <class T>
int run_plain(Name nm, const Ws<T>& W, const T* b, T* x, const Cfg& c, Out& o) {
  switch (nm) {
    case CG: cg(W, b, x, c, o); break;
    case CR: cr(W, b, x, c, o); break;
    case GCR: gcr(W, b, x, c, o); break;
    case BICG: bicg(W, b, x, c, o); break;
    case GMRES: gmres(W, b, x, c, o, nullptr); break;
    case KSKIPCG: kskipcg(W, b, x, c, o); break;
    case KSKIPBICG: kskipbicg(W, b, x, c, o); break;
    default: return KS_ERR_UNSUPPORTED;
  }
  return 0;
}

// innerMethods::innerSelect (innerMethods.cpp:71-118): a fresh inner solver
// from z = 0 with the inner parameters; in FP32 when inner_precision is 32.
template <class T>
struct Inner : Precond {
  Name nm;
  Ws<T> W;
  Cfg c;
  std::vector<T> bt, xt;
  Inner(Name nm_, const Ws<T>& w, const Cfg& c_, i64 n) : nm(nm_), W(w), c(c_), bt(n), xt(n) {}
  void apply(const double* v, double* z, Out& o) override {
    i64 n = W.A->n;
    for (i64 i = 0; i < n; ++i) { bt[i] = (T)v[i]; xt[i] = 0; }
    Out io;
    run_plain(nm, W, bt.data(), xt.data(), c, io);
    o.inner_steps += io.steps;
    o.spmv += io.spmv;
    for (i64 i = 0; i < n; ++i) z[i] = (double)xt[i];
  }
};

Cfg outer_cfg(const ks_params* p) {
  Cfg c;
  c.maxloop = p->outer_maxloop; c.eps = p->outer_eps; c.restart = p->outer_restart;
  c.kskip = p->outer_kskip; c.fix = p->outer_fix; c.compat = p->ref_compat;
  c.nchunk = p->omp_threads > 0 ? p->omp_threads : 8;
  return c;
}
Cfg inner_cfg(const ks_params* p) {
  Cfg c;
  c.maxloop = p->inner_maxloop; c.eps = p->inner_eps; c.restart = p->inner_restart;
  c.kskip = p->inner_kskip; c.fix = p->inner_fix; c.compat = p->ref_compat; c.inner = true;
  c.nchunk = p->omp_threads > 0 ? p->omp_threads : 8;
  return c;
}

}  // namespace

extern "C" {

void or_set_kskip_oob(double v) { g_kskip_oob = v; }
// k-skip CG trace: on != 0 clears and enables it, 0 disables it; the size and
// contents are read with or_kskip_trace_size / or_kskip_trace_copy.
void or_kskip_trace(int32_t on) {
  g_kskip_trace_on = on != 0;
  g_kskip_trace.clear();
}
int64_t or_kskip_trace_size(void) { return (int64_t)g_kskip_trace.size(); }
void or_kskip_trace_copy(double* out) {
  std::memcpy(out, g_kskip_trace.data(), g_kskip_trace.size() * sizeof(double));
}

void or_spmv(int64_t n, const int32_t* ptr, const int32_t* col, const double* val, const double* x, double* y) {
  Csr<double> A; A.n = n; A.ptr = ptr; A.col = col; A.val = val;
  spmv(A, x, y);
}
void or_spmv_f32(int64_t n, const int32_t* ptr, const int32_t* col, const float* val, const float* x, float* y) {
  Csr<float> A; A.n = n; A.ptr = ptr; A.col = col; A.val = val;
  spmv(A, x, y);
}
void or_transpose(int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col, const double* val,
                  int32_t* tptr, int32_t* tcol, double* tval) {
  transpose(n, nnz, ptr, col, val, tptr, tcol, tval);
}
double or_dot(int64_t n, const double* x, const double* y, int32_t nchunk) { return dot(n, x, y, nchunk); }
double or_norm2(int64_t n, const double* x, int32_t nchunk) { return norm2(n, x, nchunk); }
double or_cgs(int64_t n, int32_t k, const double* V, double* w, double* h, int32_t nchunk) {
  double hn = cgs(n, k, V, w, h, 1, nchunk);
  h[k + 1] = hn;
  return hn;
}
void or_hye(const double* h, int32_t ld, double* y, const double* e, int32_t size) { hye(h, ld, y, e, size); }
void or_kskip_basis(int64_t n, const int32_t* ptr, const int32_t* col, const double* val, int32_t k,
                    const double* r, const double* p, double* Ar, double* Ap) {
  Csr<double> A; A.n = n; A.ptr = ptr; A.col = col; A.val = val;
  kskip_basis(A, k, r, p, Ar, Ap);
}
void or_kskip_gram(int64_t n, int32_t k, const double* r, const double* p, const double* Ar, const double* Ap,
                   double* delta, double* eta, double* zeta) {
  kskip_cg_gram(n, k, r, p, Ar, Ap, delta, eta, zeta);
}
void or_sddmm(int64_t n, const int32_t* ptr, const int32_t* col, double alpha, const double* u, const double* v,
              double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    for (int32_t j = ptr[i]; j < ptr[i + 1]; ++j) out[j] = alpha * u[i] * v[col[j]];
}

// Check_error (blas.cpp:266-301) plus the plain relative residual.
void or_check(int64_t n, const int32_t* ptr, const int32_t* col, const double* val, const double* b,
              const double* x, const double* x0, double* log10_check, double* relres) {
  std::vector<double> r1(n), r0(n);
  for (int64_t i = 0; i < n; ++i) {
    double t1 = 0, t2 = 0;
    for (int32_t j = ptr[i]; j < ptr[i + 1]; ++j) {
      t1 += val[j] * x[col[j]];
      t2 += val[j] * (x0 ? x0[col[j]] : 0.0);
    }
    r1[i] = b[i] - t1;
    r0[i] = b[i] - t2;
  }
  double a = norm2(n, r1.data(), 8), c = norm2(n, r0.data(), 8), bn = norm2(n, b, 8);
  if (log10_check) *log10_check = std::log10(a / c);
  if (relres) *relres = a / bn;
}

// Full solve: outerMethods::outerSelect (outerMethods.cpp:20-82).
int32_t or_solve(const ks_params* p, int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col,
                 const double* val, const double* b, double* x, double* hist, int64_t hist_cap, ks_result* res) {
  if (res) std::memset(res, 0, sizeof(*res));
  Name on = name_of(p->outer), in = name_of(p->inner);
  bool vp = (on == VPCG || on == VPGMRES || on == VPGCR);
  if (on == NONE || on == KSKIPCR || on == VPBICG || on == VPCR) return KS_ERR_UNSUPPORTED;
  if (vp && in == NONE) return KS_ERR_BAD_INPUT;
  bool need_t = (on == BICG || on == KSKIPBICG || in == BICG || in == KSKIPBICG);
  std::vector<int32_t> tptr, tcol;
  std::vector<double> tval;
  Csr<double> A; A.n = n; A.nnz = nnz; A.ptr = ptr; A.col = col; A.val = val;
  Csr<double> AT = A;
  if (need_t) {
    tptr.resize(n + 1); tcol.resize(nnz); tval.resize(nnz);
    transpose(n, nnz, ptr, col, val, tptr.data(), tcol.data(), tval.data());
    AT.ptr = tptr.data(); AT.col = tcol.data(); AT.val = tval.data();
  }
  Ws<double> W; W.A = &A; W.AT = &AT;
  std::vector<double> x0(n, 0.0);
  if (!(p->flags & KS_X0_GIVEN)) std::fill(x, x + n, 0.0);
  for (int64_t i = 0; i < n; ++i) x0[i] = x[i];
  Out o;
  o.record = hist != nullptr;
  Cfg oc = outer_cfg(p), ic = inner_cfg(p);
  int rc = 0;
  if (!vp) {
    rc = run_plain(on, W, b, x, oc, o);
  } else {
    std::vector<float> fval, ftval;
    Csr<float> F, FT;
    Precond* M = nullptr;
    Inner<double>* m64 = nullptr;
    Inner<float>* m32 = nullptr;
    if (p->inner_precision == KS_F32) {
      fval.assign(val, val + nnz);
      F.n = n; F.nnz = nnz; F.ptr = ptr; F.col = col; F.val = fval.data();
      FT = F;
      if (need_t) {
        ftval.assign(tval.begin(), tval.end());
        FT.ptr = tptr.data(); FT.col = tcol.data(); FT.val = ftval.data();
      }
      Ws<float> WF; WF.A = &F; WF.AT = &FT;
      m32 = new Inner<float>(in, WF, ic, n);
      M = m32;
    } else {
      m64 = new Inner<double>(in, W, ic, n);
      M = m64;
    }
    if (on == VPCG) vpcg(W, b, x, oc, o, M);
    else if (on == VPGCR) vpgcr(W, b, x, oc, o, M);
    else gmres(W, b, x, oc, o, M);
    delete m32;
    delete m64;
  }
  if (rc) return rc;
  if (res) {
    res->status = o.status;
    res->iters = o.iters;
    res->steps = o.steps;
    res->inner_steps = o.inner_steps;
    res->spmv = o.spmv;
    res->relres = o.relres;
    or_check(n, ptr, col, val, b, x, x0.data(), &res->log10_check, &res->relres_true);
    int64_t nh = (int64_t)o.hist.size() / 2;
    if (hist) {
      int64_t m = std::min(nh, hist_cap);
      std::memcpy(hist, o.hist.data(), sizeof(double) * 2 * m);
    }
    res->n_hist = nh;
  }
  return o.status;
}

// Adjoint reference: A^T lambda = g with the solver named in p on the
// explicit transpose, grad_b = lambda, grad_vals = -lambda_row * x_col.
int32_t or_adjoint(const ks_params* p, int64_t n, int64_t nnz, const int32_t* ptr, const int32_t* col,
                   const double* val, const double* x, const double* g, double* grad_b, double* grad_vals,
                   ks_result* res) {
  std::vector<int32_t> tptr(n + 1), tcol(nnz);
  std::vector<double> tval(nnz);
  transpose(n, nnz, ptr, col, val, tptr.data(), tcol.data(), tval.data());
  int32_t st = or_solve(p, n, nnz, tptr.data(), tcol.data(), tval.data(), g, grad_b, nullptr, 0, res);
  if (st != 0 && st != 2) return st;
  or_sddmm(n, ptr, col, -1.0, grad_b, x, grad_vals);
  return st;
}

int32_t or_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
