// tval3_gpu_common.hpp — shared part of the C++ overload layer over the C-ABI.
//
// Not included directly: include tval3_gpu_2d.hpp or tval3_gpu_3d.hpp AFTER the
// header that defines the reference's types (mat<T>, sparse_mat<T>, geometry,
// tval3_options<T>, tval3_info<T>, tval3_exception and the format constants
// mat_row_major, mat_col_major, sparse_mat_both; the reference includes them
// from tval3_types.h at 2d/tval3_cpu_2d.cpp:1-14 and 3d/tval3_cpu_3d.cpp:1-30).
//
// The overload layer takes the reference's host containers, stages them in
// device memory through tv3_mem_*, builds a plan for A, runs tv3_solve (float
// vectors) or tv3_solve_f64 (double/double) on a process-wide context and
// copies U back.  Status codes become
// tval3_exception with the reference's "Argument error: ..." /
// "Internal error: ..." prefixes (2d/tval3_cpu_2d.cpp:696-750, 878-884).
//
// Settings read from the environment when the context is first created:
//   TVAL3_GPU_DEVICE    HIP device ordinal (default 0)
//   TVAL3_GPU_LS_POLY   1: quadratic-model Armijo search (one interpolation
//                       pass per step); default 0: the reference's
//                       round-by-round backtracking
//   TVAL3_GPU_PROFILE   1: time every launch of each solve and print the
//                       per-slot table to stdout afterwards, one line per
//                       slot in the reference's PROFILING print format
//                       (3d/tval3_cpu_3d.cpp:1337-1348); default 0
#ifndef TVAL3_GPU_COMMON_HPP_
#define TVAL3_GPU_COMMON_HPP_

#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "tval3_amd.h"

namespace tval3_gpu_detail {

struct Session {
  std::mutex mu;
  tv3_ctx *ctx = nullptr;
  int ls_poly = 0;
  int profile = 0;
};

inline Session &session() {
  static Session s;
  return s;
}

template <class Exc>
[[noreturn]] inline void fail(tv3_ctx *c, int rc) {
  const std::string what = c ? tv3_ctx_error(c) : std::string("no context");
  if (rc == TV3_ERR_ARG) throw Exc(std::string("Argument error: ") + what);
  throw Exc(std::string("Internal error: ") + what);
}

// the process-wide context (created on first use; the caller holds s.mu)
template <class Exc>
inline tv3_ctx *context(Session &s) {
  if (!s.ctx) {
    const char *dv = std::getenv("TVAL3_GPU_DEVICE");
    const char *lp = std::getenv("TVAL3_GPU_LS_POLY");
    s.ls_poly = (lp && std::atoi(lp) != 0) ? 1 : 0;
    const char *pr = std::getenv("TVAL3_GPU_PROFILE");
    s.profile = (pr && std::atoi(pr) != 0) ? 1 : 0;
    tv3_ctx *c = nullptr;
    const int rc = tv3_ctx_create(dv ? std::atoi(dv) : 0, nullptr, &c);
    if (rc) {
      const std::string why = c ? tv3_ctx_error(c) : "";
      if (c) tv3_ctx_destroy(c);
      throw Exc(std::string("Internal error: GPU context: ") + why);
    }
    s.ctx = c;
  }
  return s.ctx;
}

// device buffer owned for the duration of one call
template <class Exc>
class DevBuf {
 public:
  DevBuf(tv3_ctx *c, size_t bytes) : c_(c) {
    const int rc = tv3_mem_alloc(c, bytes, &p_);
    if (rc) fail<Exc>(c, rc);
  }
  ~DevBuf() {
    if (p_) (void)tv3_mem_free(c_, p_);
  }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  template <class T>
  T *as() const { return static_cast<T *>(p_); }
  void upload(const void *src, size_t bytes) {
    const int rc = tv3_mem_upload(c_, p_, src, bytes);
    if (rc) fail<Exc>(c_, rc);
  }
  void download(void *dst, size_t bytes) const {
    const int rc = tv3_mem_download(c_, dst, p_, bytes);
    if (rc) fail<Exc>(c_, rc);
  }

 private:
  tv3_ctx *c_;
  void *p_ = nullptr;
};

class PlanHolder {
 public:
  PlanHolder() = default;
  ~PlanHolder() {
    if (p) tv3_plan_destroy(p);
  }
  PlanHolder(const PlanHolder &) = delete;
  PlanHolder &operator=(const PlanHolder &) = delete;
  tv3_plan *p = nullptr;
};

// tval3_options<T> (18 reference fields) -> tv3_options
template <class Opts>
inline tv3_options to_c(const Opts &o, int scal_double, int ls_poly, int submat_mb) {
  tv3_options c;
  tv3_options_default(&c);
  c.mu = (double)o.mu;
  c.beta = (double)o.beta;
  c.mu0 = (double)o.mu0;
  c.beta0 = (double)o.beta0;
  c.tol = (double)o.tol;
  c.tol_inn = (double)o.tol_inn;
  c.maxit = (int)o.maxit;
  c.maxcnt = (int)o.maxcnt;
  c.gam = (double)o.gam;
  c.gamma = (double)o.gamma;
  c.rate_gam = (double)o.rate_gam;
  c.rate_cnt = (double)o.rate_cnt;
  c.c = (double)o.c;
  c.nonneg = o.nonneg ? 1 : 0;
  c.isreal = o.isreal ? 1 : 0;
  c.numInitIterations = (unsigned)o.numInitIterations;
  c.skipMulRatio = (double)o.skipMulRatio;
  c.maxRelativeChange = (double)o.maxRelativeChange;
  c.scal_double = scal_double;
  c.tv_iso = 0;
  c.dt_exact = 0;
  c.ls_exact = ls_poly ? 0 : 1;
  c.submat_mb = submat_mb > 0 ? submat_mb : 5;
  return c;
}

template <class Info>
inline void from_c(const tv3_info &i, Info &out) {
  out.secs = i.secs;
  out.outer_iters = i.outer_iters;
  out.total_iters = i.total_iters;
  out.rel_chg = i.rel_chg;
  out.rel_error = i.rel_error;
}

template <class Exc>
struct KeepAll {
  std::vector<DevBuf<Exc> *> v;
  KeepAll() = default;
  KeepAll(const KeepAll &) = delete;
  KeepAll &operator=(const KeepAll &) = delete;
  ~KeepAll() {
    for (auto *p : v) delete p;
  }
  DevBuf<Exc> *put(tv3_ctx *c, const void *src, size_t bytes) {
    v.push_back(new DevBuf<Exc>(c, bytes));
    v.back()->upload(src, bytes);
    return v.back();
  }
};

// the solve entry point for the vector type: float vectors (tv3_solve) or
// the double/double solver (tv3_solve_f64)
inline int solve_call(tv3_ctx *c, const tv3_plan *A, const tv3_options *o, int dx, int dy,
                      int dz, const float *b, float *U, const float *Ut, tv3_info *info) {
  return tv3_solve(c, A, o, dx, dy, dz, b, U, Ut, info, nullptr, 0);
}
inline int solve_call(tv3_ctx *c, const tv3_plan *A, const tv3_options *o, int dx, int dy,
                      int dz, const double *b, double *U, const double *Ut, tv3_info *info) {
  return tv3_solve_f64(c, A, o, dx, dy, dz, b, U, Ut, info, nullptr, 0);
}
// complex vectors: std::complex<T> is an interleaved (re, im) pair
inline int solve_call(tv3_ctx *c, const tv3_plan *A, const tv3_options *o, int dx, int dy,
                      int dz, const std::complex<float> *b, std::complex<float> *U,
                      const std::complex<float> *Ut, tv3_info *info) {
  return tv3_solve_c64(c, A, o, dx, dy, dz, reinterpret_cast<const float *>(b),
                       reinterpret_cast<float *>(U), reinterpret_cast<const float *>(Ut), info,
                       nullptr, 0);
}
inline int solve_call(tv3_ctx *c, const tv3_plan *A, const tv3_options *o, int dx, int dy,
                      int dz, const std::complex<double> *b, std::complex<double> *U,
                      const std::complex<double> *Ut, tv3_info *info) {
  return tv3_solve_c128(c, A, o, dx, dy, dz, reinterpret_cast<const double *>(b),
                        reinterpret_cast<double *>(U), reinterpret_cast<const double *>(Ut), info,
                        nullptr, 0);
}

// The slot table of the last solve, as the reference prints profile_info
// (name: ms per run, runs, total ms); profiling is switched off after it.
template <class Exc>
inline void print_profile(tv3_ctx *c) {
  tv3_profile_slot sl[64];
  int n = 0;
  int rc = tv3_ctx_profile_slots(c, sl, 64, &n);
  if (rc) fail<Exc>(c, rc);
  for (int i = 0; i < n && i < 64; ++i)
    std::printf("%s:\t %.2f ms (runs: %lld, total: %.1f ms)\n", sl[i].name,
                sl[i].ms / (double)sl[i].runs, sl[i].runs, sl[i].ms);
  std::fflush(stdout);
  rc = tv3_ctx_profile(c, 0);
  if (rc) fail<Exc>(c, rc);
}

// Runs one solve for host data of element type T (float, double,
// std::complex<float>, std::complex<double>).
// make_plan(ctx, PlanHolder&, KeepAll&) builds the operator plan, keeping
// its device arrays in the KeepAll; dims are U's (dim_x contiguous).
template <class Exc, class Info, class Opts, class T, class MakePlan>
inline Info solve_host(mat<T> &U, const mat<T> &b, const Opts &opts, const mat<T> &Ut,
                       int dx, int dy, int dz, int scal_double, int submat_mb,
                       MakePlan make_plan) {
  const auto t0 = std::chrono::steady_clock::now();
  Session &s = session();
  std::lock_guard<std::mutex> lock(s.mu);
  tv3_ctx *c = context<Exc>(s);
  KeepAll<Exc> keep;  // declared before the plan: released after it
  PlanHolder plan;
  make_plan(c, plan, keep);
  const size_t n = (size_t)U.len, m = (size_t)b.len;
  DevBuf<Exc> *db = keep.put(c, &b.data()[0], m * sizeof(T));
  DevBuf<Exc> dU(c, n * sizeof(T));
  const bool has_ut = Ut.len > 0 && (size_t)Ut.len == n;
  DevBuf<Exc> *dUt = has_ut ? keep.put(c, &Ut.data()[0], n * sizeof(T)) : nullptr;
  const tv3_options co = to_c(opts, scal_double, s.ls_poly, submat_mb);
  tv3_info ci{};
  if (s.profile && tv3_ctx_profile(c, 2)) fail<Exc>(c, TV3_ERR_HIP);
  const int rc = solve_call(c, plan.p, &co, dx, dy, dz, db->template as<T>(), dU.template as<T>(),
                            dUt ? dUt->template as<T>() : nullptr, &ci);
  if (rc) fail<Exc>(c, rc);
  if (s.profile) print_profile<Exc>(c);
  dU.download(&U.data()[0], n * sizeof(T));
  Info info;
  from_c(ci, info);
  info.secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return info;
}

// dense A (host, row-major float or double) -> device copy + plan
template <class Exc>
inline void dense_plan(tv3_ctx *c, PlanHolder &ph, KeepAll<Exc> &keep,
                       const float *A, int m, int n) {
  DevBuf<Exc> *a = keep.put(c, A, (size_t)m * n * sizeof(float));
  const int rc = tv3_plan_dense(c, m, n, TV3_F32, a->template as<float>(), &ph.p);
  if (rc) fail<Exc>(c, rc);
}
template <class Exc>
inline void dense_plan(tv3_ctx *c, PlanHolder &ph, KeepAll<Exc> &keep,
                       const double *A, int m, int n) {
  DevBuf<Exc> *a = keep.put(c, A, (size_t)m * n * sizeof(double));
  const int rc = tv3_plan_dense(c, m, n, TV3_F64, a->template as<double>(), &ph.p);
  if (rc) fail<Exc>(c, rc);
}
template <class Exc, class R>
inline void dense_plan(tv3_ctx *c, PlanHolder &ph, KeepAll<Exc> &keep,
                       const std::complex<R> *A, int m, int n) {
  DevBuf<Exc> *a = keep.put(c, A, (size_t)m * n * sizeof(std::complex<R>));
  const int rc = tv3_plan_dense(c, m, n, sizeof(R) == 4 ? TV3_C64 : TV3_C128, a->template as<void>(), &ph.p);
  if (rc) fail<Exc>(c, rc);
}

inline int sparse_call(tv3_ctx *c, int m, int n, int nnz, const int *rp, const int *ri,
                       const float *rv, const int *cp, const int *ci, const float *cv,
                       tv3_plan **out) {
  return tv3_plan_sparse(c, m, n, nnz, rp, ri, rv, cp, ci, cv, out);
}
inline int sparse_call(tv3_ctx *c, int m, int n, int nnz, const int *rp, const int *ri,
                       const double *rv, const int *cp, const int *ci, const double *cv,
                       tv3_plan **out) {
  return tv3_plan_sparse_f64(c, m, n, nnz, rp, ri, rv, cp, ci, cv, out);
}
inline int sparse_call(tv3_ctx *c, int m, int n, int nnz, const int *rp, const int *ri,
                       const std::complex<float> *rv, const int *cp, const int *ci,
                       const std::complex<float> *cv, tv3_plan **out) {
  return tv3_plan_sparse_c64(c, m, n, nnz, rp, ri, reinterpret_cast<const float *>(rv), cp, ci,
                             reinterpret_cast<const float *>(cv), out);
}
inline int sparse_call(tv3_ctx *c, int m, int n, int nnz, const int *rp, const int *ri,
                       const std::complex<double> *rv, const int *cp, const int *ci,
                       const std::complex<double> *cv, tv3_plan **out) {
  return tv3_plan_sparse_c128(c, m, n, nnz, rp, ri, reinterpret_cast<const double *>(rv), cp, ci,
                              reinterpret_cast<const double *>(cv), out);
}

// CSR + CSC (host, zero-based, float or double values) -> device copies + plan
template <class Exc, class SM>
inline void sparse_plan(tv3_ctx *c, PlanHolder &ph, KeepAll<Exc> &keep, const SM &A) {
  typedef typename std::decay<decltype(A.csr_val[0])>::type V;
  const int m = (int)A.dim_y, n = (int)A.dim_x;
  const int nnz = (int)A.csr_ptr[m];
  DevBuf<Exc> *rp = keep.put(c, &A.csr_ptr[0], sizeof(int) * (size_t)(m + 1));
  DevBuf<Exc> *ri = keep.put(c, &A.csr_ind[0], sizeof(int) * (size_t)nnz);
  DevBuf<Exc> *rv = keep.put(c, &A.csr_val[0], sizeof(V) * (size_t)nnz);
  DevBuf<Exc> *cp = keep.put(c, &A.csc_ptr[0], sizeof(int) * (size_t)(n + 1));
  DevBuf<Exc> *ci = keep.put(c, &A.csc_ind[0], sizeof(int) * (size_t)nnz);
  DevBuf<Exc> *cv = keep.put(c, &A.csc_val[0], sizeof(V) * (size_t)nnz);
  const int rc = sparse_call(c, m, n, nnz, rp->template as<int>(), ri->template as<int>(),
                             rv->template as<V>(), cp->template as<int>(), ci->template as<int>(),
                             cv->template as<V>(), &ph.p);
  if (rc) fail<Exc>(c, rc);
}

// geometry description (host arrays, MP-major) -> plan
template <class Exc>
inline void geometry_plan(tv3_ctx *c, PlanHolder &ph, const tv3_geometry &g) {
  const int rc = tv3_plan_geometry(c, &g, &ph.p);
  if (rc) fail<Exc>(c, rc);
}

}  // namespace tval3_gpu_detail

#endif  // TVAL3_GPU_COMMON_HPP_
