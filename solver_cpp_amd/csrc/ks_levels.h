// ks_levels.h - device-resident Krylov solvers.
//
// A Level is one solve (outer, or the inner solve of a variable-
// preconditioned outer step).  It enqueues a static schedule of gated
// kernels: enqueue_init() once, then enqueue_unit(u) for u < max_units()
// (one iteration, one Arnoldi step or one k-skip block), then
// enqueue_finish().  Stop tests run on the device; once a level has
// stopped its remaining kernels return at once.  Outer levels are driven by
// run_outer(), which enqueues batches of units and checks the host-mapped
// stop flag with one batch of look-ahead; inner levels are enqueued whole,
// gated on the outer step's stop flag as well.
#pragma once
#include <memory>

#include "ks_dev.h"

namespace ks {

// ------------------------------------------------------- stop tests -----
__device__ inline void ctl_reset(Ctl* c) {
  c->stop = 0; c->done = 0; c->status = 2; c->ksz = 0;
  c->count = 0; c->nhist = 0; c->steps = 0;
}
// relres = num / bnorm (0 when both are 0: b = 0 has the solution x = 0)
__device__ inline double relres(const Ctl* c, double num) {
  return (c->bnorm == 0.0 && num == 0.0) ? 0.0 : num / c->bnorm;
}
__device__ inline bool eps_test(Ctl* c, double idx, double num) {
  const double err = relres(c, num);
  ctl_hist(c, idx, err);
  if (err <= c->eps) { ctl_stop(c, 0); return true; }
  if (!finite(err)) { ctl_stop(c, 3); return true; }
  return false;
}
// CG family: test at the top of loop `count` (cg.cpp:184-198); the loop
// ends without a test once count > maxloop.
__device__ inline void cg_check(Ctl* c, double rr) {
  if (c->count > c->maxloop) { ctl_stop(c, 2); return; }
  eps_test(c, (double)c->count, sqrt(rr));
}
// k-skip family: test at the top of outer step nloop (kskipcg.cpp:214-227).
__device__ inline void kskip_check(Ctl* c, double rr) {
  if (c->count >= c->maxloop) { ctl_stop(c, 2); return; }
  eps_test(c, (double)(c->count + 1), sqrt(rr));
}
// GMRES family: test of step k at iteration count (gmres.cpp:278-324).
__device__ inline void gm_check(Ctl* c, int k, double num) {
  const double err = relres(c, num);
  ctl_hist(c, (double)(c->count + 1), err);
  if (c->count + 1 >= c->maxloop) { c->ksz = k; ctl_stop(c, 2); }
  else if (err <= c->eps) { c->ksz = k; ctl_stop(c, 0); }
  else if (!finite(err)) { c->ksz = k; ctl_stop(c, 3); }
}

// ------------------------------------------------------------- ops ------
// x = 0 (optional), r = (T)src, p = r (optional), acc (r,r); sets the loop.
// MODE 0: CG-like (count=1, test), 1: k-skip (count=0, test), 2: no test.
template <class T, int MODE>
struct OpCgInit {
  const double* src;
  T *x, *r, *p, *rs, *ps;
  Ctl* c;
  int zero_x, set_bnorm;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double* acc) {
    const T ri = (T)src[i];
    r[i] = ri;
    if (p) p[i] = ri;
    if (rs) rs[i] = ri;
    if (ps) ps[i] = ri;
    if (zero_x) x[i] = (T)0;
    acc[0] += (double)ri * (double)ri;
  }
  __device__ void fin(const double* s) {
    ctl_reset(c);
    if (set_bnorm) c->bnorm = sqrt(s[0]);
    c->rr = s[0];
    c->rrs[0] = s[0];
    if (MODE == 0) { c->count = 1; cg_check(c, s[0]); }
    else if (MODE == 1) { c->count = 0; kskip_check(c, s[0]); }
  }
};

struct FinAlpha {
  Ctl* c;
  __device__ void operator()(const double* s) const { c->dotv = s[0]; c->alpha = c->rr / s[0]; }
};

// x += a p; r -= a q; acc (r,r); the finisher counts the step and tests
// (VPCG and CR, whose beta comes from their own passes).
template <class T, int MODE>
struct OpCgUpdate {
  T *x, *r;
  const T *p, *q;
  Ctl* c;
  double a;
  __device__ void prep() { a = c->alpha; }
  __device__ void elem(int64_t i, double* acc) {
    x[i] = (T)(a * (double)p[i] + (double)x[i]);
    const T ri = (T)((-a) * (double)q[i] + (double)r[i]);
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
  }
  __device__ void fin(const double* s) {
    static_assert(MODE == 1, "CG takes OpCgUpdateF / OpXpbyF");
    c->count++;
    c->steps++;
    c->tsteps++;
    cg_check(c, s[0]);
  }
};

// CG's step (cg.cpp:199-240) with its two one-workgroup finishers folded
// into the kernels that follow them.  OpCgUpdateF forms alpha = rr / (p, q)
// from the SpMV's partials (FinAlpha's step) and OpXpbyF forms beta =
// rr' / rr, the new rr and the stop test from the update's partials, each in
// every workgroup with partial_totals, the finisher's own summation: the same
// bits as finisher kernels would give.  Workgroup 0 alone writes the control block.
// rr is kept by iteration parity: in iteration k both kernels read rrs[k & 1]
// and OpXpbyF's workgroup 0 writes rrs[(k + 1) & 1], so no workgroup reads a
// slot that this kernel writes.  The update's partials go to a second buffer
// (part2), as the SpMV's are still being read while the update's
// workgroups finish.  Once OpXpbyF has set the stop flag its remaining
// workgroups may skip p, which the stopped solve no longer reads.
template <class T>
struct OpCgUpdateF {
  T *x, *r;
  const T *p, *q;
  Ctl* c;
  const double* dpart;  // the SpMV's (p, q) partials
  int dg, slot;
  double a;
  __device__ void prep() {
    const double dot = partial_totals<1>(dg, dpart)[0];
    a = c->rrs[slot] / dot;
    if (blockIdx.x == 0 && threadIdx.x == 0) { c->dotv = dot; c->alpha = a; }
  }
  __device__ void elem(int64_t i, double* acc) {
    x[i] = (T)(a * (double)p[i] + (double)x[i]);
    const T ri = (T)((-a) * (double)q[i] + (double)r[i]);
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
  }
  __device__ void fin(const double*) {}
};
template <class T>
struct OpXpbyF {
  T* p;
  const T* r;
  Ctl* c;
  const double* rpart;  // the update's (r, r) partials
  int rg, slot;
  double b;
  __device__ void prep() {
    const double rr = partial_totals<1>(rg, rpart)[0];
    b = rr / c->rrs[slot];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      c->beta = b;
      c->rr = rr;
      c->rrs[slot ^ 1] = rr;
      c->count++;
      c->steps++;
      c->tsteps++;
      cg_check(c, rr);
    }
  }
  __device__ void elem(int64_t i, double*) { p[i] = (T)(b * (double)p[i] + (double)r[i]); }
  __device__ void fin(const double*) {}
};

// p = beta p + r  (blas::Scalar_axy(beta, p, r, p)), optionally twice.
template <class T>
struct OpXpby {
  T* p;
  const T* r;
  T* p2;
  const T* r2;
  const Ctl* c;
  double b;
  __device__ void prep() { b = c->beta; }
  __device__ void elem(int64_t i, double*) {
    p[i] = (T)(b * (double)p[i] + (double)r[i]);
    if (p2) p2[i] = (T)(b * (double)p2[i] + (double)r2[i]);
  }
  __device__ void fin(const double*) {}
};

// x += a p; r -= a q   (NV 0), or with acc (r,r) and the k-skip test (NV 1).
template <class T, int NV, int MODE>  // k-skip BiCG: MODE 0 inner steps, 1 the block's last
struct OpAxpy2 {
  T *x, *r;
  const T *p, *q;
  Ctl* c;
  int kp1;
  double a;
  __device__ void prep() { a = c->alpha; }
  __device__ void elem(int64_t i, double* acc) {
    x[i] = (T)(a * (double)p[i] + (double)x[i]);
    const T ri = (T)((-a) * (double)q[i] + (double)r[i]);
    r[i] = ri;
    if (NV) acc[0] += (double)ri * (double)ri;
  }
  __device__ void fin(const double* s) {
    if (NV) { c->count += kp1; c->steps += kp1; c->tsteps += kp1; kskip_check(c, s[0]); }
  }
};

// BiCG: rs -= a q2; acc0 (rs,r), acc1 (r,r), whose totals OpXpby2F forms
// (beta, rr, count++, test).
template <class T>
struct OpBicgUpdate {
  T* rs;
  const T *q2, *r;
  Ctl* c;
  double a;
  __device__ void prep() { a = c->alpha; }
  __device__ void elem(int64_t i, double* acc) {
    const T v = (T)((-a) * (double)q2[i] + (double)rs[i]);
    rs[i] = v;
    const double ri = (double)r[i];
    acc[0] += (double)v * ri;
    acc[1] += ri * ri;
  }
  __device__ void fin(const double*) {}
};

// BiCG's step (bicg.cpp:218-274) with its finishers folded forward as CG's
// (OpCgUpdateF): the x, r update forms alpha = rr / (p*, A p) from the
// SpMV's partials, and the direction update forms beta = (r*', r') / (r*, r),
// the new (r*, r) and the (r, r) stop test from OpBicgUpdate's two partial
// sums (k_finish<2>'s summation), with rr by iteration parity.
template <class T>
struct OpAxpy2F {
  T *x, *r;
  const T *p, *q;
  Ctl* c;
  const double* dpart;  // the SpMV's (p*, A p) partials
  int dg, slot;
  double a;
  __device__ void prep() {
    const double dot = partial_totals<1>(dg, dpart)[0];
    a = c->rrs[slot] / dot;
    if (blockIdx.x == 0 && threadIdx.x == 0) { c->dotv = dot; c->alpha = a; }
  }
  __device__ void elem(int64_t i, double*) {
    x[i] = (T)(a * (double)p[i] + (double)x[i]);
    r[i] = (T)((-a) * (double)q[i] + (double)r[i]);
  }
  __device__ void fin(const double*) {}
};
template <class T>
struct OpXpby2F {
  T* p;
  const T* r;
  T* p2;
  const T* r2;
  Ctl* c;
  const double* upart;  // OpBicgUpdate's ((r*, r), (r, r)) partials
  int ug, slot;
  double b;
  __device__ void prep() {
    const double* s = partial_totals<2>(ug, upart);
    const double rz = s[0], rr = s[1];
    b = rz / c->rrs[slot];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      c->beta = b;
      c->rr = rz;
      c->rrs[slot ^ 1] = rz;
      c->count++;
      c->steps++;
      c->tsteps++;
      cg_check(c, rr);
    }
  }
  __device__ void elem(int64_t i, double*) {
    p[i] = (T)(b * (double)p[i] + (double)r[i]);
    p2[i] = (T)(b * (double)p2[i] + (double)r2[i]);
  }
  __device__ void fin(const double*) {}
};

// k-skip: x += a p; r -= a Av; p = b p + r; last step of a block adds the
// (r,r) test (NV 1).
template <class T, int NV>
struct OpAxpy3 {
  T *x, *r, *p;
  const T* q;
  Ctl* c;
  int kp1;
  double a, b;
  __device__ void prep() { a = c->alpha; b = c->beta; }
  __device__ void elem(int64_t i, double* acc) {
    const double pi = (double)p[i];
    x[i] = (T)(a * pi + (double)x[i]);
    const T ri = (T)((-a) * (double)q[i] + (double)r[i]);
    r[i] = ri;
    p[i] = (T)(b * pi + (double)ri);
    if (NV) acc[0] += (double)ri * (double)ri;
  }
  __device__ void fin(const double* s) {
    if (NV) { c->count += kp1; c->steps += kp1; c->tsteps += kp1; kskip_check(c, s[0]); }
  }
};

// VPCG: p = z, acc0 (r,z), acc1 (r,r); test at loop 1.
struct OpVpInit {
  double* p;
  const double *z, *r;
  Ctl* c;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double* acc) {
    const double zi = z[i], ri = r[i];
    p[i] = zi;
    acc[0] += ri * zi;
    acc[1] += ri * ri;
  }
  __device__ void fin(const double* s) {
    c->rr = s[0];
    c->count = 1;
    cg_check(c, s[1]);
  }
};
// VPCG beta: (r',z')/(r,z) (vpcg.cpp:753) or flexible -alpha (z',Ap)/(r,z).
struct OpVpBeta {
  const double *r, *z, *q;
  Ctl* c;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double* acc) {
    const double zi = z[i];
    acc[0] += r[i] * zi;
    acc[1] += zi * q[i];
  }
  __device__ void fin(const double* s) {
    c->beta = (c->compat & KS_COMPAT_VPCG_FR) ? s[0] / c->rr : -c->alpha * s[1] / c->rr;
    c->rr = s[0];
  }
};

// out = in / vscale (blas::Scalar_x_div_a)
template <class T>
struct OpScale {
  const T* in;
  T* out;
  const Ctl* c;
  double vs;
  __device__ void prep() { vs = c->vscale; }
  __device__ void elem(int64_t i, double*) { out[i] = (T)((double)in[i] / vs); }
  __device__ void fin(const double*) {}
};

// GMRES cycle start: r = (T)src (first cycle, x0 = 0) and the norm.  x is
// not written: its finisher sets xzero, and the assembly takes x as zero.
template <class T>
struct OpGmFirst {
  const double* src;
  T *r, *x;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double* acc) {
    const T ri = (T)src[i];
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
  }
};
// xz: x is the zero the solve started from (the first cycle with x0 = 0)
struct FinGmCycle {
  Ctl* c;
  int first, set_bnorm, xz;
  __device__ void operator()(const double* s) const {
    const double t = sqrt(s[0]);
    if (first) ctl_reset(c);
    c->xzero = xz;
    if (set_bnorm) c->bnorm = t;
    for (int i = 0; i < c->R; ++i) c->e[i] = 0.0;
    c->e[0] = t;
    c->vscale = t;
    c->vn[0] = t;
    c->vr[0] = 1.0 / t;
    gm_check(c, 0, t);
  }
};
template <class T>
struct OpGmFirstFin : OpGmFirst<T> {
  FinGmCycle f;
  __device__ void fin(const double* s) { f(s); }
};
// OpScale of the level above (v = w / vscale, FP64) fused with an inner
// GMRES cycle start on v: r = (T)v and ||r||^2 in one pass (x = 0 as in
// OpGmFirst) - the same values as the two passes.
template <class T>
struct OpGmFirstScaled {
  const double* w;
  const Ctl* sc;  // the level above: its vscale
  double* v;
  T *r, *x;
  double vs;
  FinGmCycle f;
  __device__ void prep() { vs = sc->vscale; }
  __device__ void elem(int64_t i, double* acc) {
    const double vi = w[i] / vs;
    v[i] = vi;
    const T ri = (T)vi;
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
  }
  __device__ void fin(const double* s) { f(s); }
};
struct FinGmNorm {
  Ctl* c;
  int k;
  __device__ void operator()(const double* s) const { gm_norm_step(c, k, sqrt(s[0])); }
};

// Residual/bnorm helper ops for the API entry points.
struct OpNorm2 {
  const double* x;
  double* out;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double* acc) { acc[0] += x[i] * x[i]; }
  __device__ void fin(const double* s) { *out = sqrt(s[0]); }
};
struct FinSetBnorm {
  Ctl* c;
  __device__ void operator()(const double* s) const { c->bnorm = sqrt(s[0]); }
};
struct OpSetBnorm {
  const double* b;
  Ctl* c;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double* acc) { acc[0] += b[i] * b[i]; }
  __device__ void fin(const double* s) { c->bnorm = sqrt(s[0]); }
};

// k-skip CG scalar recurrences for inner step s (kskipcg.cpp:266-288).
struct OpKskipScalar {
  Ctl* c;
  int s;
  __device__ void run() {
    const int k = c->kskip;
    double* delta = c->delta;
    double* eta = c->eta;
    double* zeta = c->zeta;
    const double alpha = c->gamma / zeta[0];
    const double beta = alpha * zeta[1] / zeta[0] - 1.0;
    if (c->fix == 1) {
      c->gamma = beta * c->gamma;
    } else if (c->fix == 2) {
      const double t0 = c->gamma - alpha * eta[0];
      const double t1 = eta[0] - alpha * zeta[1];
      c->gamma = t0 - alpha * t1;
    }
    const int jm = 2 * k - 2 * s;
    if (c->compat & KS_COMPAT_KSKIP_AS_WRITTEN) {
      for (int j = 0; j < jm; ++j) {
        delta[j] = delta[j] - 2 * alpha * eta[j + 1] + alpha * alpha * eta[j + 2];
        const double eta_old = eta[j];
        eta[j] = delta[j] + beta * zeta[j + 1] - alpha * beta * zeta[j + 1];
        zeta[j] = eta[j + 1] + beta * eta_old + beta * beta * zeta[j] - alpha * beta * zeta[j + 1];
      }
    } else {
      for (int j = 0; j < jm; ++j) {
        const double dn = delta[j] - 2 * alpha * eta[j + 1] + alpha * alpha * zeta[j + 2];
        const double en = dn + beta * eta[j] - alpha * beta * zeta[j + 1];
        const double zn = en + beta * eta[j] - alpha * beta * zeta[j + 1] + beta * beta * zeta[j];
        delta[j] = dn;
        eta[j] = en;
        zeta[j] = zn;
      }
    }
    c->alpha = alpha;
    c->beta = beta;
  }
};
// k-skip BiCG scalar recurrences (kskipbicg.cpp:298-314).
struct OpKskipBicgScalar {
  Ctl* c;
  int s;
  __device__ void run() {
    const int k = c->kskip;
    double* th = c->delta;
    double* eta = c->eta;
    double* rho = c->rho;
    double* phi = c->phi;
    const double alpha = c->gamma / phi[0];
    const double beta = 1.0 - (eta[0] + rho[0] - alpha * phi[1]) / phi[0];
    c->gamma = c->gamma * beta;
    const int jm = 2 * k - 2 * s;
    for (int j = 0; j < jm; ++j) {
      th[j] = th[j] - alpha * eta[j + 1] - alpha * rho[j + 1] + alpha * alpha * phi[j + 2];
      const double t = th[j] - alpha * beta * phi[j + 1];
      eta[j] = t + beta * eta[j];
      rho[j] = t + beta * rho[j];
      phi[j] = eta[j] + rho[j] - th[j] + beta * beta * phi[j];
    }
    c->alpha = alpha;
    c->beta = beta;
  }
};

// ----------------------------------------------------------- Level -----
struct LevelCfg {
  int maxloop = 0;
  double eps = 0;
  int restart = 0;
  int kskip = 0;
  int fix = 0;
  unsigned compat = 0;
  int64_t hist_cap = 0;
};

// n rounded up to a multiple of 64 elements (256-byte aligned columns)
constexpr int64_t pad64(int64_t n) { return (n + 63) & ~(int64_t)63; }

struct Level {
  ks_ctx* c = nullptr;
  int64_t n = 0;
  LevelCfg cfg;
  DBuf<Ctl> dctl;
  Ctl hctl{};
  DBuf<double> arr;   // H, e, cs, sn, y, delta, eta, zeta, rho, phi
  DBuf<double> hist;
  const int* outer_stop = nullptr;  // set for inner levels
  bool x0given = false;
  virtual ~Level() {}

  Ctl* ctl() const { return dctl.p; }
  const int* stop_flag() const { return &dctl.p->stop; }
  Gate gstep() const { return Gate{outer_stop, &dctl.p->stop}; }
  Gate gdone() const { return Gate{outer_stop, &dctl.p->done}; }
  Gate gouter() const { return Gate{outer_stop, nullptr}; }

  // Doubles of device memory setup_ctl(R, k, mem) takes from `mem`: the
  // small arrays, then the control block on a 256-byte boundary.
  static constexpr int64_t ctl_doubles(int R, int k) {
    return pad64(small_doubles(R, k)) + pad64((int64_t)((sizeof(Ctl) + 7) / 8));
  }
  static constexpr int64_t small_doubles(int R, int k) {
    const int64_t nH = R > 0 ? (int64_t)(R + 1) * R : 0;
    const int64_t nsmall = R > 0 ? 6 * (int64_t)R : 0;
    return nH + nsmall + 5 * (int64_t)(2 * k + 3) + 8;
  }
  // mem (optional): device memory of ctl_doubles(R, k) doubles that holds the
  // small arrays and the control block instead of allocations (cleared here
  // with a fill kernel); no history with it.
  void setup_ctl(int R, int k, double* mem = nullptr) {
    hctl = Ctl{};
    hctl.status = 2;
    hctl.maxloop = cfg.maxloop;
    hctl.eps = cfg.eps;
    hctl.R = R;
    hctl.ld = R > 0 ? R : 1;
    hctl.kskip = k;
    hctl.fix = cfg.fix;
    hctl.compat = cfg.compat;
    const int64_t nH = R > 0 ? (int64_t)(R + 1) * R : 0;
    const int64_t narr = small_doubles(R, k);
    if (mem) {
      KS_CHECK(cfg.hist_cap == 0, KS_ERR_BAD_INPUT, "no residual history without allocations");
      arr.borrow(mem, (size_t)narr);
      k_fill<double><<<grid_map(narr), kBlock, 0, c->stream>>>(narr, arr.p, 0.0);
      KS_HIPCK(hipGetLastError());
    } else {
      arr.alloc((size_t)narr);
      KS_HIPCK(hipMemsetAsync(arr.p, 0, arr.n * sizeof(double), c->stream));
    }
    double* q = arr.p;
    if (R > 0) {
      hctl.H = q; q += nH;
      hctl.e = q; q += R;
      hctl.cs = q; q += R;
      hctl.sn = q; q += R;
      hctl.y = q; q += R;
      hctl.vn = q; q += R;
      hctl.vr = q; q += R;
    }
    hctl.delta = q; q += 2 * k + 3;
    hctl.eta = q; q += 2 * k + 3;
    hctl.zeta = q; q += 2 * k + 3;
    hctl.rho = q; q += 2 * k + 3;
    hctl.phi = q; q += 2 * k + 3;
    if (cfg.hist_cap > 0) {
      hist.alloc((size_t)(2 * cfg.hist_cap));
      hctl.hist = hist.p;
      hctl.hist_cap = cfg.hist_cap;
    }
    if (mem)
      dctl.borrow((Ctl*)(mem + pad64(narr)), 1);
    else
      dctl.alloc(1);
    upload_ctl();
  }
  void upload_ctl() { KS_HIPCK(hipMemcpyAsync(dctl.p, &hctl, sizeof(Ctl), hipMemcpyHostToDevice, c->stream)); }

  // b (FP64, device) is the right-hand side; x_out (FP64, device) receives
  // the solution.  Inner levels always start from x = 0.
  virtual void enqueue_init(const double* b, double* x_out) = 0;
  // Solve with the right-hand side v = w / sc->vscale, which this call also
  // writes to v (the level above's basis column).  By default the scaling is
  // a pass of its own; a level may fuse it into its first pass.
  virtual void enqueue_init_scaled(const double* w, const Ctl* sc, double* v, double* x_out) {
    launch_map<0>(c, n, OpScale<double>{w, v, sc, 0.0}, gouter(), KS_K_AXPY, 16.0 * n);
    enqueue_init(v, x_out);
  }
  virtual void enqueue_unit(int64_t u) = 0;
  virtual int64_t max_units() const = 0;
  virtual void enqueue_finish() {}
  virtual int64_t reported_iters(const Ctl& h) const { return h.count; }
  virtual double unit_seconds_estimate() const = 0;
};

// ------------------------------------------------------------- CG ------
template <class T>
struct CgLevel : Level {
  MatView<T> A;
  DBuf<T> r, p, q, xown;
  T* x = nullptr;
  double* xout = nullptr;
  // Doubles of device memory setup(..., mem) takes: r, p and q (each padded
  // to 64 elements), then the small arrays and the control block.
  static constexpr int64_t scratch_doubles(int64_t n) { return 3 * pad64(n) + ctl_doubles(0, 0); }
  // mem (optional, FP64 only): device memory of scratch_doubles(n) doubles
  // that holds every vector and the control state instead of allocations
  void setup(ks_ctx* cx, const MatView<T>& a, const LevelCfg& cf, double* mem = nullptr) {
    c = cx; A = a; n = a.n; cfg = cf;
    if (mem) {
      KS_CHECK(sizeof(T) == 8, KS_ERR_UNSUPPORTED, "borrowed CG state is FP64 only");
      const int64_t ld = pad64(n);
      r.borrow((T*)mem, n);
      p.borrow((T*)(mem + ld), n);
      q.borrow((T*)(mem + 2 * ld), n);
      setup_ctl(0, 0, mem + 3 * ld);
      return;
    }
    r.alloc(n); p.alloc(n); q.alloc(n);
    if (sizeof(T) != 8) xown.alloc(n);
    setup_ctl(0, 0);
  }
  void enqueue_init(const double* b, double* x_out) override {
    xout = x_out;
    x = sizeof(T) == 8 ? (T*)x_out : xown.p;
    const double* src = b;
    if (x0given) {  // outer FP64 only: bnorm, then r = b - A x0
      launch_map<1>(c, n, OpSetBnorm{b, ctl()}, gouter(), KS_K_REDUCE, 8.0 * n);
      launch_spmv<T, 1>(c, A, x, EpiResid<T, FinNone>{b, r.p, FinNone{}}, gouter(), 8.0 * n);
      src = (const double*)r.p;
    }
    OpCgInit<T, 0> op{src, x, r.p, p.p, nullptr, nullptr, ctl(), !x0given, !x0given};
    launch_map<1>(c, n, op, gouter(), KS_K_REDUCE, (x0given ? 24.0 : 32.0) * n);
  }
  // q = A p with (p, q) partials, then the update and the direction, which
  // take the totals themselves (OpCgUpdateF, OpXpbyF): three launches per
  // iteration, no finisher
  void enqueue_unit(int64_t u) override {
    const int slot = (int)(u & 1);
    const int gd = launch_spmv<T, 1>(c, A, p.p, EpiDot<T, FinAlpha>{p.p, q.p, FinAlpha{ctl()}}, gstep(),
                                     sizeof(T) * (double)n, false);
    const int gu = launch_map_partials<1>(c, n, OpCgUpdateF<T>{x, r.p, p.p, q.p, ctl(), c->part.p, gd, slot, 0.0},
                                          gstep(), KS_K_AXPY, 6.0 * sizeof(T) * n, RedArgs{c->part2.p});
    launch_map<0>(c, n, OpXpbyF<T>{p.p, r.p, ctl(), c->part2.p, gu, slot, 0.0}, gstep(), KS_K_AXPY,
                  3.0 * sizeof(T) * n);
  }
  int64_t max_units() const override { return cfg.maxloop > 0 ? cfg.maxloop : 0; }
  void enqueue_finish() override {
    if (sizeof(T) != 8) {
      k_convert<T, double><<<grid_map(n), kBlock, 0, c->stream>>>(n, x, xout, gouter());
      KS_HIPCK(hipGetLastError());
    }
  }
  double unit_seconds_estimate() const override {
    return std::max(3 * 4e-6, (spmv_bytes<T>(n, A.nnz) + 10.0 * sizeof(T) * n) / 4e12);
  }
};

// ------------------------------------------------------------- CR ------
// cr::solve (cr.cpp:152-280): alpha = (r,s)/(q,q), s = A r,
// beta = (r',s')/(r,s), p = beta p + r, q = beta q + s.  (q,q) is summed in
// the pass that forms q; (r,s) in the SpMV that forms s.
struct FinCrBeta {
  Ctl* c;
  __device__ void operator()(const double* s) const {
    c->beta = s[0] / c->rz;
    c->rz = s[0];
  }
};
// after q = A p (init): s = q, acc0 (r,s), acc1 (q,q), acc2 (r,r); test at loop 1
template <class T>
struct OpCrInit {
  const T *r, *q;
  T* s;
  Ctl* c;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double* acc) {
    const double ri = (double)r[i], qi = (double)q[i];
    s[i] = q[i];
    acc[0] += ri * qi;
    acc[1] += qi * qi;
    acc[2] += ri * ri;
  }
  __device__ void fin(const double* x) {
    c->rz = x[0];
    c->alpha = x[0] / x[1];
    c->count = 1;
    cg_check(c, x[2]);
  }
};
// p = beta p + r; q = beta q + s; acc (q,q) -> alpha of the next iteration
template <class T>
struct OpCrXpby {
  T *p, *q;
  const T *r, *s;
  Ctl* c;
  double b;
  __device__ void prep() { b = c->beta; }
  __device__ void elem(int64_t i, double* acc) {
    p[i] = (T)(b * (double)p[i] + (double)r[i]);
    const T qi = (T)(b * (double)q[i] + (double)s[i]);
    q[i] = qi;
    acc[0] += (double)qi * (double)qi;
  }
  __device__ void fin(const double* x) { c->alpha = c->rz / x[0]; }
};

template <class T>
struct CrLevel : Level {
  MatView<T> A;
  DBuf<T> r, p, q, s, xown;
  T* x = nullptr;
  double* xout = nullptr;
  void setup(ks_ctx* cx, const MatView<T>& a, const LevelCfg& cf) {
    c = cx; A = a; n = a.n; cfg = cf;
    r.alloc(n); p.alloc(n); q.alloc(n); s.alloc(n);
    if (sizeof(T) != 8) xown.alloc(n);
    setup_ctl(0, 0);
  }
  void enqueue_init(const double* b, double* x_out) override {
    xout = x_out;
    x = sizeof(T) == 8 ? (T*)x_out : xown.p;
    const double* src = b;
    if (x0given) {
      launch_map<1>(c, n, OpSetBnorm{b, ctl()}, gouter(), KS_K_REDUCE, 8.0 * n);
      launch_spmv<T, 1>(c, A, x, EpiResid<T, FinNone>{b, r.p, FinNone{}}, gouter(), 8.0 * n);
      src = (const double*)r.p;
    }
    OpCgInit<T, 2> op{src, x, r.p, p.p, nullptr, nullptr, ctl(), !x0given, !x0given};
    launch_map<1>(c, n, op, gouter(), KS_K_REDUCE, 32.0 * n);
    launch_spmv<T, 0>(c, A, p.p, EpiStore<T>{q.p}, gouter());
    launch_map<3>(c, n, OpCrInit<T>{r.p, q.p, s.p, ctl()}, gouter(), KS_K_REDUCE, 3.0 * sizeof(T) * n);
  }
  void enqueue_unit(int64_t) override {
    launch_map<1>(c, n, OpCgUpdate<T, 1>{x, r.p, p.p, q.p, ctl(), 0.0}, gstep(), KS_K_AXPY, 6.0 * sizeof(T) * n);
    launch_spmv<T, 1>(c, A, r.p, EpiDot<T, FinCrBeta>{r.p, s.p, FinCrBeta{ctl()}}, gstep(), sizeof(T) * (double)n);
    launch_map<1>(c, n, OpCrXpby<T>{p.p, q.p, r.p, s.p, ctl(), 0.0}, gstep(), KS_K_AXPY, 6.0 * sizeof(T) * n);
  }
  int64_t max_units() const override { return cfg.maxloop > 0 ? cfg.maxloop : 0; }
  void enqueue_finish() override {
    if (sizeof(T) != 8) {
      k_convert<T, double><<<grid_map(n), kBlock, 0, c->stream>>>(n, x, xout, gouter());
      KS_HIPCK(hipGetLastError());
    }
  }
  double unit_seconds_estimate() const override {
    return std::max(3 * 4e-6, (spmv_bytes<T>(n, A.nnz) + 12.0 * sizeof(T) * n) / 4e12);
  }
};

// ----------------------------------------------------------- BiCG ------
template <class T>
struct BicgLevel : Level {
  MatView<T> A, AT;
  DBuf<T> r, rs, p, ps, q, xown;
  T* x = nullptr;
  double* xout = nullptr;
  void setup(ks_ctx* cx, const MatView<T>& a, const MatView<T>& at, const LevelCfg& cf) {
    c = cx; A = a; AT = at; n = a.n; cfg = cf;
    r.alloc(n); rs.alloc(n); p.alloc(n); ps.alloc(n); q.alloc(n);
    if (sizeof(T) != 8) xown.alloc(n);
    setup_ctl(0, 0);
  }
  void enqueue_init(const double* b, double* x_out) override {
    xout = x_out;
    x = sizeof(T) == 8 ? (T*)x_out : xown.p;
    const double* src = b;
    if (x0given) {
      launch_map<1>(c, n, OpSetBnorm{b, ctl()}, gouter(), KS_K_REDUCE, 8.0 * n);
      launch_spmv<T, 1>(c, A, x, EpiResid<T, FinNone>{b, r.p, FinNone{}}, gouter(), 8.0 * n);
      src = (const double*)r.p;
    }
    OpCgInit<T, 0> op{src, x, r.p, p.p, rs.p, ps.p, ctl(), !x0given, !x0given};
    launch_map<1>(c, n, op, gouter(), KS_K_REDUCE, 48.0 * n);
  }
  // no finisher: OpAxpy2F and OpXpby2F take the totals they need
  void enqueue_unit(int64_t u) override {
    const int slot = (int)(u & 1);
    const int gd = launch_spmv<T, 1>(c, A, p.p, EpiDot<T, FinAlpha>{ps.p, q.p, FinAlpha{ctl()}}, gstep(),
                                     sizeof(T) * (double)n, false);
    launch_map<0>(c, n, OpAxpy2F<T>{x, r.p, p.p, q.p, ctl(), c->part.p, gd, slot, 0.0}, gstep(), KS_K_AXPY,
                  6.0 * sizeof(T) * n);
    launch_spmv<T, 0>(c, AT, ps.p, EpiStore<T>{q.p}, gstep());
    const int gu = launch_map_partials<2>(c, n, OpBicgUpdate<T>{rs.p, q.p, r.p, ctl(), 0.0}, gstep(), KS_K_AXPY,
                                          4.0 * sizeof(T) * n, red(c));
    launch_map<0>(c, n, OpXpby2F<T>{p.p, r.p, ps.p, rs.p, ctl(), c->part.p, gu, slot, 0.0}, gstep(), KS_K_AXPY,
                  6.0 * sizeof(T) * n);
  }
  int64_t max_units() const override { return cfg.maxloop > 0 ? cfg.maxloop : 0; }
  void enqueue_finish() override {
    if (sizeof(T) != 8) {
      k_convert<T, double><<<grid_map(n), kBlock, 0, c->stream>>>(n, x, xout, gouter());
      KS_HIPCK(hipGetLastError());
    }
  }
  double unit_seconds_estimate() const override {
    return std::max(5 * 4e-6, (2 * spmv_bytes<T>(n, A.nnz) + 20.0 * sizeof(T) * n) / 4e12);
  }
};

// --------------------------------------------------- projection helper --
// h[i] = (w, V_i), i <= k, in launches of up to 32 columns.
// 16-byte loads need 16-byte aligned columns (base and ldv).
template <class T>
bool vec_ok(const T* V, int64_t ldv, const T* w) {
  return ((uintptr_t)V % 16 == 0) && ((uintptr_t)w % 16 == 0) && (ldv % Vec16<T>::N == 0);
}

template <class T, int GV>
void proj_launch(ks_ctx* c, int64_t n, const T* Vi, int64_t ldv, int m, const T* w, double* hi, int ldh, Gate g,
                 const double* vn) {
  const double bytes = (double)sizeof(T) * n * (m + 1);
  int G;
  if (vec_ok(Vi, ldv, w)) {
    G = grid_map(c, k_proj<T, GV, true>, (n + Vec16<T>::N - 1) / Vec16<T>::N);
    klaunch(c, KS_K_CGS, k_proj<T, GV, true>, G, 0, n, Vi, ldv, m, w, g, red(c), stat(c, KS_K_CGS, bytes));
  } else {
    G = grid_map(c, k_proj<T, GV, false>, n);
    klaunch(c, KS_K_CGS, k_proj<T, GV, false>, G, 0, n, Vi, ldv, m, w, g, red(c), stat(c, KS_K_CGS, bytes));
  }
  launch_finish<GV>(c, G, FinProj{hi, ldh, m, vn}, g);
}

// vn (device, nullable): norms of unnormalised basis columns.
template <class T>
void launch_proj(ks_ctx* c, int64_t n, const T* V, int64_t ldv, int nv, const T* w, double* hcol, int ldh,
                 Gate g, const double* vn = nullptr) {
  for (int i0 = 0; i0 < nv; i0 += 32) {
    const int m = std::min(32, nv - i0);
    const T* Vi = V + (int64_t)i0 * ldv;
    double* hi = hcol + (int64_t)i0 * ldh;
    const double* vi = vn ? vn + i0 : nullptr;
    if (m <= 4) proj_launch<T, 4>(c, n, Vi, ldv, m, w, hi, ldh, g, vi);
    else if (m <= 8) proj_launch<T, 8>(c, n, Vi, ldv, m, w, hi, ldh, g, vi);
    else if (m <= 16) proj_launch<T, 16>(c, n, Vi, ldv, m, w, hi, ldh, g, vi);
    else proj_launch<T, 32>(c, n, Vi, ldv, m, w, hi, ldh, g, vi);
  }
}

template <class Fin>
struct FinWrap {
  Fin f;
  __device__ void fin(const double* s) const { f(s); }
};

// The projection finisher, sized for nv columns.
template <class F>
void launch_proj_finish(ks_ctx* c, int G, int nv, const F& f, Gate g) {
  if (nv <= 4) launch_finish<4>(c, G, f, g);
  else if (nv <= 8) launch_finish<8>(c, G, f, g);
  else if (nv <= 16) launch_finish<16>(c, G, f, g);
  else if (nv <= 32) launch_finish<32>(c, G, f, g);
  else launch_finish<64>(c, G, f, g);
}

// The deferred norm step of GMRES step pk (its ||v_{pk+1}|| from the ng
// partials np[], formed as the fused SpMV formed it), then the projection
// finisher of step pk + 1 unless that norm step stopped the solve.
struct FinNormProj {
  FinProj p;
  Ctl* c;
  int pk;
  const double* np;
  int ng;
  double h;
  __device__ void pre() { h = norm_from_partials(np, ng); }
  __device__ void fin(const double* s) {
    gm_norm_step(c, pk, h);
    if (c->stop) return;
    p.fin(s);
  }
};

// Whether launch_spmv_proj fuses (and so can also take a deferred norm).
template <class T>
bool spmv_proj_fused(const MatView<T>& A, const T* V, int64_t ldv, int nv, const T* w) {
  return A.kind == 0 && A.wg_tiles && nv <= kMaxNV && vec_ok(V, ldv, w);
}

// Deferred norm step of GMRES step pk: its update kernel left ng partials of
// ||v_{pk+1}||^2 in c->part2 and no finisher ran (pk < 0: none pending).
struct PendNorm {
  Ctl* c = nullptr;
  int pk = -1;
  int ng = 0;
};

// w = (A vk) * r (r null: w = A vk) and h_i = (w, V_i) / vn_i (vn null: no
// division), i < nv.  On workgroup tiles one launch does both (the
// CSR-stream kernel's projection epilogue), then the projection finisher;
// otherwise the SpMV and a separate projection.  With pn pending, the
// previous step's norm finisher runs first: folded into the fused launch
// (which then takes its scale from the partials, not from r), or as its own
// launch.
template <class T>
void launch_spmv_proj(ks_ctx* c, const MatView<T>& A, const T* vk, T* w, const double* r, const T* V, int64_t ldv,
                      int nv, double* hcol, int ldh, Gate g, const double* vn, PendNorm pn = PendNorm{}) {
  if (spmv_proj_fused(A, V, ldv, nv, (const T*)w)) {
    const int cls = sizeof(T) == 8 ? KS_K_SPMV64 : KS_K_SPMV32;
    const double bytes = spmv_bytes<T>(A.n, A.nnz, A.ib) + (double)sizeof(T) * A.n * nv;
    EpiDivProj<T> e{};
    e.y = w; e.r = r; e.V = V; e.ldv = ldv; e.nv = nv;
    if (pn.pk >= 0) {
      e.np = c->part2.p; e.ng = pn.ng;
    }
    // the int32 kernel's grid in every index format (see launch_stream)
    const int G = grid_rows(c, k_spmv_stream<T, 1, 0, EpiDivProj<T>>, A.ntiles);
    launch_stream<T, 1, 0, EpiDivProj<T>>(c, cls, A, G, vk, vk, e, g, bytes);
    if (pn.pk >= 0) launch_proj_finish(c, G, nv, FinNormProj{FinProj{hcol, ldh, nv, vn}, pn.c, pn.pk,
                                                             c->part2.p, pn.ng, 0.0}, g);
    else launch_proj_finish(c, G, nv, FinProj{hcol, ldh, nv, vn}, g);
  } else {
    if (pn.pk >= 0) launch_finish<1>(c, pn.ng, FinWrap<FinGmNorm>{FinGmNorm{pn.c, pn.pk}}, g, c->part2.p);
    if (r) launch_spmv<T, 0>(c, A, vk, EpiDiv<T>{w, r}, g);
    else launch_spmv<T, 0>(c, A, vk, EpiStore<T>{w}, g);
    launch_proj<T>(c, A.n, V, ldv, nv, w, hcol, ldh, g, vn);
  }
}

// out = w - V h (out may be w); vn as in launch_proj.
// defer: the ||out||^2 partials go to c->part2 and no finisher is launched
// (the next fused SpMV takes them, see EpiDivProj); returns the grid size.
template <class T, class Fin>
int launch_cgs_update(ks_ctx* c, int64_t n, const T* V, int64_t ldv, int nv, const T* w, T* out,
                      const double* hcol, int ldh, Fin f, Gate g, const double* vn = nullptr, bool defer = false) {
  const double bytes = (double)sizeof(T) * n * (nv + 2);
  const size_t lds = sizeof(double) * std::max(nv, 1);
  const RedArgs ra = defer ? RedArgs{c->part2.p} : red(c);
  int G;
  if (vec_ok(V, ldv, w) && vec_ok(V, ldv, (const T*)out)) {
    G = grid_map(c, k_cgs_update<T, true>, (n + Vec16<T>::N - 1) / Vec16<T>::N, lds);
    klaunch(c, KS_K_CGS, k_cgs_update<T, true>, G, lds, n, V, ldv, nv, w, out, hcol, ldh, vn, g, ra,
            stat(c, KS_K_CGS, bytes));
  } else {
    G = grid_map(c, k_cgs_update<T, false>, n, lds);
    klaunch(c, KS_K_CGS, k_cgs_update<T, false>, G, lds, n, V, ldv, nv, w, out, hcol, ldh, vn, g, ra,
            stat(c, KS_K_CGS, bytes));
  }
  if (!defer) launch_finish<1>(c, G, FinWrap<Fin>{f}, g);
  return G;
}


// ---------------------------------------------------------- GMRES ------
// gmres::solve (gmres.cpp:232-423); with `inner` set it is vpgmres::solve
// (vpgmres.cpp:746-944): z_k = M(v_k) into Z[k], w = A z_k, x from Z.
template <class T>
struct GmresLevel : Level {
  MatView<T> A;
  int R = 0;
  int64_t ldv = 0;
  DBuf<T> V, w, r, xown;
  PendNorm pend;  // plain GMRES: a norm step deferred into the next step's SpMV
  DBuf<double> Z;  // VP outer only (T = double)
  T* x = nullptr;
  double* xout = nullptr;
  const double* b = nullptr;
  // enqueue_init_scaled: b = sw / ssc->vscale, formed by the first cycle start
  const double* sw = nullptr;
  const Ctl* ssc = nullptr;
  std::unique_ptr<Level> inner;
  void setup(ks_ctx* cx, const MatView<T>& a, const LevelCfg& cf, std::unique_ptr<Level> in = nullptr) {
    c = cx; A = a; n = a.n; cfg = cf; R = cf.restart;
    KS_CHECK(R >= 2, KS_ERR_BAD_INPUT, "GMRES restart (-R/-r) must be >= 2");
    ldv = pad64(n);
    V.alloc((size_t)ldv * R);
    w.alloc(n);
    if (in) r.alloc(n);  // plain GMRES forms r in V_0
    if (sizeof(T) != 8) xown.alloc(n);
    setup_ctl(R, 0);
    inner = std::move(in);
    if (inner) {
      Z.alloc((size_t)ldv * R);
      inner->outer_stop = stop_flag_ptr();
    }
  }
  const int* stop_flag_ptr() { return &dctl.p->stop; }
  int64_t cycles() const {
    if (cfg.maxloop <= 0) return 0;
    return (int64_t)(cfg.maxloop - 1) / (R - 1) + 1;
  }
  // The test of step k runs in unit k-1 (the Givens finisher) except at a
  // cycle start, so unit maxloop-1 does no work unless it starts a cycle.
  int64_t max_units() const override {
    if (cfg.maxloop <= 0) return 0;
    const int64_t last = cfg.maxloop - 1;
    const int64_t u = last % (R - 1) == 0 ? last + 1 : last;
    return std::min<int64_t>(u, cycles() * (R - 1));
  }
  void enqueue_init(const double* bb, double* x_out) override {
    b = bb;
    xout = x_out;
    x = sizeof(T) == 8 ? (T*)x_out : xown.p;
    pend = PendNorm{};
    sw = nullptr;
    ssc = nullptr;
    if (x0given) launch_map<1>(c, n, OpSetBnorm{b, ctl()}, gouter(), KS_K_REDUCE, 8.0 * n);
  }
  // the scaling rides on the first cycle start (x0 = 0), which reads the
  // right-hand side anyway
  void enqueue_init_scaled(const double* w, const Ctl* sc, double* v, double* x_out) override {
    if (x0given) {
      Level::enqueue_init_scaled(w, sc, v, x_out);
      return;
    }
    enqueue_init(v, x_out);
    sw = w;
    ssc = sc;
  }
  // Plain GMRES keeps its basis unnormalised: V_k holds w_k with norm
  // vn[k], the SpMV of step k divides its rows by vn[k], and the projection,
  // the update and the assembly take the norms into their coefficients, so no
  // separate v = w / ||w|| pass runs.  VPGMRES normalises V explicitly: its
  // inner solve takes v_k as a right-hand side.
  void cycle_start(int64_t cyc) {
    T* r0 = inner ? r.p : V.p;
    if (cyc == 0 && !x0given && sw) {
      OpGmFirstScaled<T> op{sw, ssc, const_cast<double*>(b), r0, x, 0.0, FinGmCycle{ctl(), 1, 1, 1}};
      launch_map<1>(c, n, op, gouter(), KS_K_REDUCE, (16.0 + 2.0 * sizeof(T)) * n);
      sw = nullptr;
    } else if (cyc == 0 && !x0given) {
      OpGmFirstFin<T> op;
      op.src = b; op.r = r0; op.x = x; op.f = FinGmCycle{ctl(), 1, 1, 1};
      launch_map<1>(c, n, op, gouter(), KS_K_REDUCE, (8.0 + 2.0 * sizeof(T)) * n);
    } else {
      Gate g = cyc == 0 ? gouter() : gstep();
      launch_spmv<T, 1>(c, A, x, EpiResid<T, FinGmCycle>{b, r0, FinGmCycle{ctl(), cyc == 0 ? 1 : 0, 0, 0}}, g,
                        8.0 * n);
    }
    // VPGMRES: v_0 = r / ||r|| is formed by the inner solve's first pass
    // (enqueue_init_scaled in enqueue_unit)
  }
  void enqueue_unit(int64_t u) override {
    const int64_t cyc = u / (R - 1);
    const int k = (int)(u % (R - 1));
    if (k == 0) cycle_start(cyc);
    T* vk = V.p + (int64_t)k * ldv;
    T* vk1 = V.p + (int64_t)(k + 1) * ldv;
    Ctl* h = ctl();
    if (inner) {
      // v_k = r / ||r|| (k = 0) or w / ||w|| (w from step k - 1), formed by
      // the inner solve's first pass on it
      double* zk = Z.p + (int64_t)k * ldv;
      inner->enqueue_init_scaled(k == 0 ? (const double*)r.p : (const double*)w.p, h, (double*)vk, zk);
      for (int64_t iu = 0; iu < inner->max_units(); ++iu) inner->enqueue_unit(iu);
      inner->enqueue_finish();
      launch_spmv_proj<T>(c, A, (const T*)zk, w.p, nullptr, V.p, ldv, k + 1, hctl.H + k, hctl.ld, gstep(), nullptr);
      launch_cgs_update<T>(c, n, V.p, ldv, k + 1, w.p, w.p, hctl.H + k, hctl.ld, FinGmNorm{h, k}, gstep());
    } else {
      launch_spmv_proj<T>(c, A, vk, w.p, hctl.vr + k, V.p, ldv, k + 1, hctl.H + k, hctl.ld, gstep(), hctl.vn,
                          pend);
      pend = PendNorm{};
      // the norm step moves into the next step's fused SpMV when that step
      // is in this cycle and fuses (flushed by enqueue_finish otherwise)
      const bool defer = k < R - 2 && spmv_proj_fused(A, (const T*)V.p, ldv, k + 2, (const T*)w.p);
      const int G = launch_cgs_update<T>(c, n, V.p, ldv, k + 1, w.p, vk1, hctl.H + k, hctl.ld, FinGmNorm{h, k},
                                         gstep(), hctl.vn, defer);
      if (defer) pend = PendNorm{h, k, G};
    }
    if (k == R - 2) cycle_end();
  }
  void cycle_end() {
    Ctl* h = ctl();
    k_hye<<<1, 1, 0, c->stream>>>(h, gdone());
    KS_HIPCK(hipGetLastError());
    const int last = (cfg.compat & KS_COMPAT_LAST_TERM) ? 1 : 0;
    // an FP32 level's assembly also writes its FP64 output (xout), so no
    // conversion pass follows: every change of x is an assembly, the last
    // one included (enqueue_finish's cycle_end when the solve stopped at a
    // cycle start)
    double* xw = sizeof(T) != 8 ? xout : nullptr;
    const double bytes = (double)sizeof(T) * n * 2 + (inner ? 8.0 : (double)sizeof(T)) * n * (R - 1) +
                         (xw ? 8.0 * n : 0.0);
    const size_t lds = sizeof(double) * R;
    if (inner)
      klaunch(c, KS_K_CGS, k_assemble<T, double>, grid_map(c, k_assemble<T, double>, n, lds), lds, n, Z.p, ldv, x,
              (double*)nullptr, h, last, 0, gdone(), stat(c, KS_K_CGS, bytes));
    else
      klaunch(c, KS_K_CGS, k_assemble<T, T>, grid_map(c, k_assemble<T, T>, n, lds), lds, n, V.p, ldv, x, xw, h,
              last, 1, gdone(), stat(c, KS_K_CGS, bytes));
    k_mark_done<<<1, 1, 0, c->stream>>>(h, gdone());
    KS_HIPCK(hipGetLastError());
  }
  void enqueue_finish() override {
    if (pend.pk >= 0) {  // the last enqueued step's norm, when no later step took it
      launch_finish<1>(c, pend.ng, FinWrap<FinGmNorm>{FinGmNorm{pend.c, pend.pk}}, gstep(), c->part2.p);
      pend = PendNorm{};
    }
    cycle_end();  // no-op unless the solve stopped in a cycle not yet assembled
  }
  int64_t reported_iters(const Ctl& hc) const override { return hc.count + 1; }
  double unit_seconds_estimate() const override {
    double s = (spmv_bytes<T>(n, A.nnz) + (double)sizeof(T) * n * (R + 8)) / 4e12 + 5 * 4e-6;
    if (inner) s += inner->max_units() * inner->unit_seconds_estimate();
    return s;
  }
};

// ------------------------------------------------------------ GCR ------
// gcr::solve (gcr.cpp:178-313); with `inner` it is vpgcr::solve
// (vpgcr.cpp:693-860): P_0 = M(r) and each new direction starts from
// z = M(r).  loop starts at 1 (gcr.cpp:41); the test order is eps first,
// then the iteration limit (gcr.cpp:226-234).  Ctl arrays: e = qq, y = (Av,q_i).
template <class T>
struct OpGcrDots {  // acc0 (q,q), acc1 (r,q), acc2 (r,r)
  const T *q, *r;
  Ctl* c;
  int k;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double* acc) {
    const double qi = (double)q[i], ri = (double)r[i];
    acc[0] += qi * qi;
    acc[1] += ri * qi;
    acc[2] += ri * ri;
  }
  __device__ void fin(const double* s) {
    const double err = relres(c, sqrt(s[2]));
    ctl_hist(c, (double)c->count, err);
    if (err <= c->eps) { ctl_stop(c, 0); return; }
    if (!finite(err)) { ctl_stop(c, 3); return; }
    if (c->count >= c->maxloop) { ctl_stop(c, 2); return; }
    c->count++;
    c->steps++;
    c->tsteps++;
    c->e[k] = s[0];
    c->alpha = s[1] / s[0];
  }
};
template <class T>
struct OpGcrX {  // x += a P_k; r -= a Q_k (unless the cycle's last step)
  T *x, *r;
  const T *pk, *qk;
  const Ctl* c;
  int upd_r;
  double a;
  __device__ void prep() { a = c->alpha; }
  __device__ void elem(int64_t i, double*) {
    x[i] = (T)(a * (double)pk[i] + (double)x[i]);
    if (upd_r) r[i] = (T)((-a) * (double)qk[i] + (double)r[i]);
  }
  __device__ void fin(const double*) {}
};
// cycle start after the first: the while (loop < maxloop) test (gcr.cpp:188)
struct OpGcrWhile {
  Ctl* c;
  __device__ void run() {
    if (c->count >= c->maxloop) ctl_stop(c, 2);
  }
};
template <class T>
struct OpCopy {
  const T* a;
  T* b;
  __device__ void prep() {}
  __device__ void elem(int64_t i, double*) { b[i] = a[i]; }
  __device__ void fin(const double*) {}
};

template <class T>
struct GcrLevel : Level {
  MatView<T> A;
  int R = 0;
  int64_t ldv = 0;
  DBuf<T> P, Q, r, Av, z, xown;
  T* x = nullptr;
  double* xout = nullptr;
  const double* b = nullptr;
  std::unique_ptr<Level> inner;  // VPGCR (T = double)
  void setup(ks_ctx* cx, const MatView<T>& a, const LevelCfg& cf, std::unique_ptr<Level> in = nullptr) {
    c = cx; A = a; n = a.n; cfg = cf; R = cf.restart;
    KS_CHECK(R >= 1, KS_ERR_BAD_INPUT, "GCR restart (-R/-r) must be >= 1");
    ldv = pad64(n);
    P.alloc((size_t)ldv * R);
    Q.alloc((size_t)ldv * R);
    r.alloc(n); Av.alloc(n);
    if (sizeof(T) != 8) xown.alloc(n);
    setup_ctl(R, 0);
    inner = std::move(in);
    if (inner) {
      z.alloc(n);
      inner->outer_stop = &dctl.p->stop;
    }
  }
  int64_t cycles() const {
    if (cfg.maxloop <= 1) return 1;
    return (int64_t)(cfg.maxloop - 1 + R - 1) / R + 1;
  }
  int64_t max_units() const override { return cycles() * R; }
  void run_inner(const double* v, double* out) {
    inner->enqueue_init(v, out);
    for (int64_t u = 0; u < inner->max_units(); ++u) inner->enqueue_unit(u);
    inner->enqueue_finish();
  }
  void enqueue_init(const double* bb, double* x_out) override {
    b = bb;
    xout = x_out;
    x = sizeof(T) == 8 ? (T*)x_out : xown.p;
    if (x0given) launch_map<1>(c, n, OpSetBnorm{b, ctl()}, gouter(), KS_K_REDUCE, 8.0 * n);
    else {
      // x = 0, r = b, bnorm = ||b||; count (loop) = 1
      OpCgInit<T, 2> op{b, x, r.p, nullptr, nullptr, nullptr, ctl(), 1, 1};
      launch_map<1>(c, n, op, gouter(), KS_K_REDUCE, (8.0 + 2.0 * sizeof(T)) * n);
    }
    launch_scalar(c, OpSetCount{ctl(), 1}, gouter());
  }
  struct OpSetCount {
    Ctl* c;
    long long v;
    __device__ void run() { c->count = v; }
  };
  void cycle_start(int64_t cyc) {
    const Gate g = gstep();
    launch_scalar(c, OpGcrWhile{ctl()}, g);
    if (cyc > 0 || x0given)
      launch_spmv<T, 0>(c, A, x, EpiResid<T, FinNone>{b, r.p, FinNone{}}, g, 8.0 * n);
    T* p0 = P.p;
    if (inner) {
      run_inner((const double*)r.p, (double*)z.p);
      launch_map<0>(c, n, OpCopy<T>{z.p, p0}, g, KS_K_AXPY, 2.0 * sizeof(T) * n);
    } else {
      launch_map<0>(c, n, OpCopy<T>{r.p, p0}, g, KS_K_AXPY, 2.0 * sizeof(T) * n);
    }
    launch_spmv<T, 0>(c, A, p0, EpiStore<T>{Q.p}, g);
  }
  void enqueue_unit(int64_t u) override {
    const int64_t cyc = u / R;
    const int k = (int)(u % R);
    if (k == 0) cycle_start(cyc);
    const Gate g = gstep();
    T* pk = P.p + (int64_t)k * ldv;
    T* qk = Q.p + (int64_t)k * ldv;
    launch_map<3>(c, n, OpGcrDots<T>{qk, r.p, ctl(), k}, g, KS_K_REDUCE, 2.0 * sizeof(T) * n);
    const int last = k == R - 1;
    launch_map<0>(c, n, OpGcrX<T>{x, r.p, pk, qk, ctl(), !last, 0.0}, g, KS_K_AXPY,
                  (last ? 3.0 : 6.0) * sizeof(T) * n);
    if (last) return;
    const T* src = r.p;
    if (inner) {
      run_inner((const double*)r.p, (double*)z.p);
      src = z.p;
    }
    launch_spmv<T, 0>(c, A, src, EpiStore<T>{Av.p}, g);
    launch_proj<T>(c, n, Q.p, ldv, k + 1, Av.p, hctl.y, 1, g);
    const double bytes = (double)sizeof(T) * n * (2.0 * (k + 1) + 4.0);
    const size_t lds = sizeof(double) * (k + 1);
    klaunch(c, KS_K_CGS, k_gcr_update<T>, grid_map(c, k_gcr_update<T>, n, lds), lds, n, P.p, Q.p, ldv, k + 1, src,
            Av.p, hctl.y, hctl.e, P.p + (int64_t)(k + 1) * ldv, Q.p + (int64_t)(k + 1) * ldv, g,
            stat(c, KS_K_CGS, bytes));
  }
  void enqueue_finish() override {
    if (sizeof(T) != 8) {
      k_convert<T, double><<<grid_map(n), kBlock, 0, c->stream>>>(n, x, xout, gouter());
      KS_HIPCK(hipGetLastError());
    }
  }
  double unit_seconds_estimate() const override {
    double s = (spmv_bytes<T>(n, A.nnz) + (double)sizeof(T) * n * (2 * R + 8)) / 4e12 + 7 * 4e-6;
    if (inner) s += inner->max_units() * inner->unit_seconds_estimate();
    return s;
  }
};

// ------------------------------------------------------- k-skip CG -----
template <class T, bool BICG>
struct KskipLevel : Level {
  MatView<T> A, AT;
  int K = 0;
  int64_t ldv = 0;
  DBuf<double> r, p, rs, ps, q, Ar, Ap;
  double* x = nullptr;
  void setup(ks_ctx* cx, const MatView<T>& a, const MatView<T>& at, const LevelCfg& cf) {
    static_assert(sizeof(T) == 8, "k-skip levels run in FP64");
    c = cx; A = a; AT = at; n = a.n; cfg = cf; K = cf.kskip;
    KS_CHECK(K >= 1 && K <= 7, KS_ERR_UNSUPPORTED, "k-skip depth must be in [1, 7]");
    ldv = pad64(n);
    r.alloc(n); p.alloc(n); q.alloc(n);
    if (BICG) { rs.alloc(n); ps.alloc(n); }
    Ar.alloc((size_t)ldv * (2 * K + 1));
    Ap.alloc((size_t)ldv * (2 * K + 2));
    setup_ctl(0, K);
  }
  void enqueue_init(const double* b, double* x_out) override {
    x = x_out;
    const double* src = b;
    if (x0given) {
      launch_map<1>(c, n, OpSetBnorm{b, ctl()}, gouter(), KS_K_REDUCE, 8.0 * n);
      launch_spmv<T, 1>(c, A, x, EpiResid<T, FinNone>{b, r.p, FinNone{}}, gouter(), 8.0 * n);
      src = r.p;
    }
    OpCgInit<T, 1> op{src, x, r.p, p.p, BICG ? rs.p : nullptr, BICG ? ps.p : nullptr, ctl(), !x0given, !x0given};
    launch_map<1>(c, n, op, gouter(), KS_K_REDUCE, (BICG ? 48.0 : 32.0) * n);
  }
  int64_t max_units() const override { return cfg.maxloop > 0 ? (cfg.maxloop + K) / (K + 1) : 0; }
  template <int KK>
  void gram() {
    using Op = OpGram<KK, BICG>;
    Op op{ldv, r.p, p.p, BICG ? rs.p : nullptr, BICG ? ps.p : nullptr, Ar.p, Ap.p, ctl()};
    const double bytes = 8.0 * n * (BICG ? (4 * KK + 7) : (4 * KK + 5));
    launch_map<Op::NV>(c, n, op, gstep(), KS_K_BASIS, bytes);
  }
  void enqueue_unit(int64_t) override {
    // matrix-powers basis (cuda::Kskip_cg_bicg_base, cudaFunction.cu:984-1026)
    const Gate g = gstep();
    for (int ii = 0; ii < 2 * K + 2; ++ii) {
      const double* sr = ii == 0 ? r.p : Ar.p + (int64_t)(ii - 1) * ldv;
      const double* sp = ii == 0 ? p.p : Ap.p + (int64_t)(ii - 1) * ldv;
      const bool need_r = BICG ? ii < 2 * K + 1 : ii < 2 * K;  // CG never reads Ar[2K]
      if (need_r)
        launch_spmv2<T>(c, A, sr, sp, Ar.p + (int64_t)ii * ldv, Ap.p + (int64_t)ii * ldv, g);
      else
        launch_spmv<T, 0>(c, A, sp, EpiStore<T>{Ap.p + (int64_t)ii * ldv}, g);
    }
    switch (K) {
      case 1: gram<1>(); break;
      case 2: gram<2>(); break;
      case 3: gram<3>(); break;
      case 4: gram<4>(); break;
      case 5: gram<5>(); break;
      case 6: gram<6>(); break;
      default: gram<7>(); break;
    }
    for (int s = 0; s <= K; ++s) {
      if (BICG) launch_scalar(c, OpKskipBicgScalar{ctl(), s}, g);
      else launch_scalar(c, OpKskipScalar{ctl(), s}, g);
      launch_spmv<T, 0>(c, A, p.p, EpiStore<T>{q.p}, g);
      if (!BICG) {
        if (s < K)
          launch_map<0>(c, n, OpAxpy3<T, 0>{x, r.p, p.p, q.p, ctl(), K + 1, 0, 0}, g, KS_K_AXPY, 56.0 * n);
        else
          launch_map<1>(c, n, OpAxpy3<T, 1>{x, r.p, p.p, q.p, ctl(), K + 1, 0, 0}, g, KS_K_AXPY, 56.0 * n);
      } else {
        if (s < K)
          launch_map<0>(c, n, OpAxpy2<T, 0, 0>{x, r.p, p.p, q.p, ctl(), K + 1, 0}, g, KS_K_AXPY, 48.0 * n);
        else
          launch_map<1>(c, n, OpAxpy2<T, 1, 1>{x, r.p, p.p, q.p, ctl(), K + 1, 0}, g, KS_K_AXPY, 48.0 * n);
        launch_spmv<T, 0>(c, AT, ps.p, EpiStore<T>{q.p}, g);
        launch_map<0>(c, n, OpRsUpdate{rs.p, q.p, ctl(), 0}, g, KS_K_AXPY, 24.0 * n);
        launch_map<0>(c, n, OpXpby<T>{p.p, r.p, ps.p, rs.p, ctl(), 0.0}, g, KS_K_AXPY, 48.0 * n);
      }
    }
  }
  // k-skip BiCG: rs -= alpha A^T ps
  struct OpRsUpdate {
    double* rs;
    const double* q;
    const Ctl* c;
    double a;
    __device__ void prep() { a = c->alpha; }
    __device__ void elem(int64_t i, double*) { rs[i] = (-a) * q[i] + rs[i]; }
    __device__ void fin(const double*) {}
  };
  int64_t reported_iters(const Ctl& hc) const override { return BICG ? hc.count + 1 : hc.count - K + 1; }
  double unit_seconds_estimate() const override {
    return (double)(5 * K + 8) * 4e-6 + ((4 * K + 3 + K + 1) * spmv_bytes<T>(n, A.nnz)) / 4e12;
  }
};

// ------------------------------------------------------------ VPCG -----
// vpcg::solve (vpcg.cpp:659-760) with z = M(r) from z = 0 each time.
struct VpcgLevel : Level {
  MatView<double> A;
  DBuf<double> r, p, q, z;
  double* x = nullptr;
  std::unique_ptr<Level> inner;
  void setup(ks_ctx* cx, const MatView<double>& a, const LevelCfg& cf, std::unique_ptr<Level> in) {
    c = cx; A = a; n = a.n; cfg = cf;
    r.alloc(n); p.alloc(n); q.alloc(n); z.alloc(n);
    inner = std::move(in);
    setup_ctl(0, 0);
    inner->outer_stop = &dctl.p->stop;
  }
  void run_inner(const double* v) {
    inner->enqueue_init(v, z.p);
    for (int64_t u = 0; u < inner->max_units(); ++u) inner->enqueue_unit(u);
    inner->enqueue_finish();
  }
  void enqueue_init(const double* b, double* x_out) override {
    x = x_out;
    const double* src = b;
    if (x0given) {
      launch_map<1>(c, n, OpSetBnorm{b, ctl()}, gouter(), KS_K_REDUCE, 8.0 * n);
      launch_spmv<double, 1>(c, A, x, EpiResid<double, FinNone>{b, r.p, FinNone{}}, gouter(), 8.0 * n);
      src = r.p;
    }
    // r = b, x = 0, bnorm; the test of loop 1 comes after z = M(r)
    OpCgInit<double, 2> op{src, x, r.p, nullptr, nullptr, nullptr, ctl(), !x0given, !x0given};
    launch_map<1>(c, n, op, gouter(), KS_K_REDUCE, 24.0 * n);
    run_inner(r.p);
    launch_map<2>(c, n, OpVpInit{p.p, z.p, r.p, ctl()}, gstep(), KS_K_REDUCE, 24.0 * n);
  }
  void enqueue_unit(int64_t) override {
    launch_spmv<double, 1>(c, A, p.p, EpiDot<double, FinAlpha>{p.p, q.p, FinAlpha{ctl()}}, gstep(), 8.0 * n);
    launch_map<1>(c, n, OpCgUpdate<double, 1>{x, r.p, p.p, q.p, ctl(), 0.0}, gstep(), KS_K_AXPY, 48.0 * n);
    run_inner(r.p);
    launch_map<2>(c, n, OpVpBeta{r.p, z.p, q.p, ctl()}, gstep(), KS_K_REDUCE, 24.0 * n);
    launch_map<0>(c, n, OpXpby<double>{p.p, z.p, nullptr, nullptr, ctl(), 0.0}, gstep(), KS_K_AXPY, 24.0 * n);
  }
  int64_t max_units() const override { return cfg.maxloop > 0 ? cfg.maxloop : 0; }
  double unit_seconds_estimate() const override {
    return (spmv_bytes<double>(n, A.nnz) + 150.0 * n) / 4e12 + 5 * 4e-6 +
           inner->max_units() * inner->unit_seconds_estimate();
  }
};

}  // namespace ks
