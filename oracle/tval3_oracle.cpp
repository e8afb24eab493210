/*
 * tval3_oracle.cpp — CPU oracle for the TVAL3 inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, the smoke check and the CPU-baseline leg of
 * the benchmark do, and there only as the checker.
 *
 * It restates, step by step, the reference solver of modk/tval3-cpu:
 *   scaleb        2d/tval3_cpu_2d.cpp:228-266   (3d:311-360)
 *   D, Dt         2d/tval3_cpu_2d.cpp:268-293   (3d:365-425, incl. the 3-D
 *                 Dt y-term that is not the adjoint of D for dim_z >= 2)
 *   get_g_mult    2d/tval3_cpu_2d.cpp:298-394   (3d:433-529)
 *   get_g         2d/tval3_cpu_2d.cpp:396-423   (3d:535-576)
 *   get_tau       2d/tval3_cpu_2d.cpp:425-453   (3d:581-619)
 *   nonneg        2d/tval3_cpu_2d.cpp:455-465
 *   descend       2d/tval3_cpu_2d.cpp:481-493
 *   update_g      2d/tval3_cpu_2d.cpp:495-524
 *   min_u         2d/tval3_cpu_2d.cpp:526-576   (Armijo test on |f|, 5-strike
 *                 fallback)
 *   get_gradient  2d/tval3_cpu_2d.cpp:578-585
 *   shrinkage     2d/tval3_cpu_2d.cpp:587-611   (3d:834-871)
 *   update_W      2d/tval3_cpu_2d.cpp:651-671
 *   update_mlp    2d/tval3_cpu_2d.cpp:673-694
 *   driver        2d/tval3_cpu_2d.cpp:752-886   (3d:1172-1378)
 *   gemv          2d/mat_mult.h:15-91, 3d/mat_mult.h:148-595 (Bresenham walk
 *                 restated from its recurrence, 3d/mat_mult.h:206-359)
 *
 * A volume with dim_z == 1 takes the 2-D code path (the reference's 2-D
 * driver); dim_z >= 2 takes the 3-D one.
 *
 * Arithmetic conventions (the reference leaves them to MKL and the compiler):
 *   - dot products and norms accumulate exact FP64 products in FP64 and round
 *     once to float (cblas_sdot / cblas_snrm2 return float);
 *   - matrix-vector products accumulate in FP64 and round once to float;
 *     geometry rows accumulate the float products x * scale_factor;
 *   - "x + a*y" vector updates (vec_add, cblas_axpy) are single-rounding FMAs
 *     in the precision the reference computes them in; cblas_axpy and
 *     cblas_scal take a float alpha (2d/blas_wrap.h:9-45);
 *   - scalar expressions are evaluated in the reference's types and order
 *     with no contraction (build with -ffp-contract=off);
 *   - the shrinkage sum lam1 accumulates in FP64 and rounds to T_scal.
 * Reductions use fixed 4096-element chunks summed in chunk order, so results
 * do not depend on the OpenMP thread count (tv3o_set_sum_chunk changes the
 * chunk length: the tests use it to see how far a different FP64 summation
 * order moves a double/double solve).
 *
 * How each part is pinned (tests/test_oracle.py, tests/test_walk_closed_form.py):
 *   Bresenham walk   golden walks from an independent transcription of the
 *                    recurrence, walk invariants, the integer closed form,
 *                    nnz counts of the ring geometries
 *   gemv             FP64 brute force (dense, CSR/CSC, walk-built matrix)
 *   D, Dt            closed form; adjoint identity; the 3-D y-term defect
 *                    equals its closed form
 *   shrinkage        soft-threshold closed forms and limits
 *   scaleb           closed form (maxit = 0 returns the scaled A^T b)
 *   get_tau, descend, get_gradient, update_W, min_u (Armijo rule and the
 *   5-strike fallback), update_mlp, continuation, nonneg
 *                    an FP64 model of the augmented Lagrangian: first steps
 *                    are exact line minimisers along grad/beta, later steps
 *                    are Barzilai-Borwein, accepted steps satisfy the
 *                    nonmonotone Armijo rule and rejected ones do not, f
 *                    equals the Lagrangian, multipliers follow the ALM update
 *   complex instantiations (Solver<std::complex<R>, R>)
 *                    products vs numpy (A^H by the adjoint identity); the
 *                    complex shrinkage closed form (both 3-D Wz forms); an
 *                    FP64 model of the complex augmented Lagrangian (first
 *                    step, BB step); exact covariance under b -> i b; phase
 *                    covariance and the real solve on real data
 *                    (tests/test_oracle_complex.py)
 *   ls_exact = 0     same steps as the round-by-round search (f(alpha) is
 *                    an exact quadratic)
 *   layers > 1       bit-identical to the single-layer solve with the
 *                    block-diagonal operator I (x) A (dense and sparse), and
 *                    independent of submat_mb (two plain products, 2d:365)
 *   3-D steps (z terms of get_tau, lam2 / lam4, shrinkage, Dt, update_mlp)
 *                    the same FP64 model of the 3-D augmented Lagrangian
 *                    with the exact Dt (dt_exact = 1): first-step line
 *                    minimiser, BB step, Armijo acceptance, f = L, the
 *                    multiplier step and the 5-strike fallback
 *   A^T skipping (get_g_mult, 2d/tval3_cpu_2d.cpp:300-357)
 *                    maxRelativeChange = 0 gives the skipMulRatio = 0 solve
 *                    bit for bit; maxRelativeChange = +inf gives the A^T
 *                    product count of the counter schedule (periodic
 *                    multiply / skip patterns for ratios 1/2, 1/4 and 2, a
 *                    reset at every forced product)
 *   planted mistakes (tv3o_plant) each fail one of these checks
 */
#include "../include/tval3_amd.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef long long i64;
// chunk length of the FP64 reductions (4096; tests change it only to measure
// how much a different summation order moves a double/double solve)
i64 g_chunk = 4096;
// Planted mistakes (tv3o_plant; 0 = none, the only setting outside the
// tests that show each one fails a pinning check): 1 first-step tau without
// the z part of |Dd|^2, 2 lam2 without |Vz|^2, 3 lam4 without <sz, Vz>,
// 4 no z multiplier step, 5 exact Dt without its z terms, 6 BB step without
// (mu/beta) dg, 7 Armijo without its slope term, 8 no gam decay in the
// 5-strike fallback, 9 no U reset in the fallback, 10 the skip counter not
// raised by skipMulRatio, 11 forced products not resetting the skip
// counter, 12 lam1 without the z shrinkage, 13 delta step with the wrong
// sign, 14 first-step tau without the (mu/beta) |Ad|^2 term.
int g_plant = 0;
// products A x / A^T x issued by the solver since tv3o_reset_counts
i64 g_nfwd = 0, g_nbwd = 0;

/* ---------------------------------------------------------------------- */
/* reductions                                                              */
/* ---------------------------------------------------------------------- */
template <class F>
double chunked_sum(i64 n, F term) {
  const i64 kChunk = g_chunk;
  const i64 nc = (n + kChunk - 1) / kChunk;
  std::vector<double> part(nc > 0 ? nc : 1, 0.0);
#pragma omp parallel for schedule(static)
  for (i64 c = 0; c < nc; ++c) {
    const i64 lo = c * kChunk, hi = std::min(n, lo + kChunk);
    double s = 0.0;
    for (i64 i = lo; i < hi; ++i) s += term(i);
    part[c] = s;
  }
  double s = 0.0;
  for (i64 c = 0; c < nc; ++c) s += part[c];
  return s;
}

/* Vector element types: float, double, std::complex<float>,
 * std::complex<double> (the reference's T_comp).  R is the real type of a
 * component. */
template <class T>
struct VT {
  typedef T R;
  static const bool cplx = false;
};
template <class R_>
struct VT<std::complex<R_> > {
  typedef R_ R;
  static const bool cplx = true;
};

/* Terms of the FP64 dot products.  Real: the exact product.  Complex
 * (cblas_?dotc, conj(a) b): Re = a.re b.re + a.im b.im and Im = a.re b.im -
 * a.im b.re, each one FP64 rounding of two FP64 products. */
inline double dre(float a, float b) { return (double)a * (double)b; }
inline double dre(double a, double b) { return a * b; }
template <class R>
double dre(const std::complex<R> &a, const std::complex<R> &b) {
  return (double)a.real() * (double)b.real() + (double)a.imag() * (double)b.imag();
}
template <class R>
double dim_(const std::complex<R> &a, const std::complex<R> &b) {
  return (double)a.real() * (double)b.imag() - (double)a.imag() * (double)b.real();
}

template <class T>
double ddot(i64 n, const T *a, const T *b) {
  return chunked_sum(n, [&](i64 i) { return dre(a[i], b[i]); });
}
/* cblas_?dot / cblas_?nrm2 (complex: the real part of cblas_?dotc, which is
 * all the solver uses but for one BB quotient): result in the component
 * type */
template <class T>
typename VT<T>::R sdot(i64 n, const T *a, const T *b) {
  return (typename VT<T>::R)ddot(n, a, b);
}
template <class T>
typename VT<T>::R sdot_im(i64 n, const T *a, const T *b) {
  return (typename VT<T>::R)chunked_sum(n, [&](i64 i) { return dim_(a[i], b[i]); });
}
template <class T>
typename VT<T>::R snrm2(i64 n, const T *a) {
  return (typename VT<T>::R)std::sqrt(ddot(n, a, a));
}

/* |z| of a complex value: FP64 square root of the FP64 sum of squares,
 * rounded once; a value on an axis gives its one component's magnitude */
inline float modulus(float a) { return std::abs(a); }
inline double modulus(double a) { return std::abs(a); }
template <class R>
R modulus(const std::complex<R> &z) {
  if (z.imag() == 0) return std::abs(z.real());
  if (z.real() == 0) return std::abs(z.imag());
  const double a = z.real(), b = z.imag();
  return (R)std::sqrt(a * a + b * b);
}

/* real scalar times a vector element, per component */
template <class R>
R mul_r(R s, R x) { return s * x; }
template <class R>
std::complex<R> mul_r(R s, const std::complex<R> &x) {
  return std::complex<R>(s * x.real(), s * x.imag());
}
/* fma(alpha, y, x) per component in precision TS, rounded to the component */
template <class TS, class R>
R fma_s(TS alpha, R y, R x) { return (R)std::fma(alpha, (TS)y, (TS)x); }
template <class TS, class R>
std::complex<R> fma_s(TS alpha, const std::complex<R> &y, const std::complex<R> &x) {
  return std::complex<R>((R)std::fma(alpha, (TS)y.real(), (TS)x.real()),
                         (R)std::fma(alpha, (TS)y.imag(), (TS)x.imag()));
}

/* z = x + alpha*y with one rounding, computed in precision TS (complex: per
 * component, as the reference's complex vec_add, 2d/tval3_cpu_2d.cpp:210-226) */
template <class T, class TS>
void vec_add(i64 n, const T *x, TS alpha, const T *y, T *z) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) z[i] = fma_s(alpha, y[i], x[i]);
}
/* z = x - y (vec_add with alpha = -1: exact negation, one rounding) */
template <class T>
void vec_sub(i64 n, const T *x, const T *y, T *z) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) z[i] = x[i] - y[i];
}
/* y += alpha*x, alpha a real value in the component type (cblas_saxpy /
 * cblas_daxpy; the complex solver's alphas are real values, applied per
 * component) */
template <class T>
void axpy(i64 n, typename VT<T>::R alpha, const T *x, T *y) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) y[i] = fma_s(alpha, x[i], y[i]);
}
template <class T>
void scal(i64 n, typename VT<T>::R alpha, T *x) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) x[i] = mul_r(alpha, x[i]);
}

/* ---------------------------------------------------------------------- */
/* measurement operators                                                   */
/* ---------------------------------------------------------------------- */
struct Op {
  int kind;  // 0 dense, 1 sparse, 2 geometry
  i64 m, n;
  // dense
  const float *A;
  // sparse (borrowed)
  const int *csr_ptr, *csr_ind, *csc_ptr, *csc_ind;
  const float *csr_val, *csc_val;
  // complex entries (the complex solver's mat<complex<T>> / sparse_mat
  // values), re and im interleaved, as doubles (complex<float> entries widen
  // exactly): dense m*n, sparse nnz
  const double *Az;
  const double *csr_valz, *csc_valz;
  i64 nnz;
  // geometry: walks in reference loop order
  float scale;
  std::vector<int> walk_ray;      // output row of each walk
  std::vector<i64> walk_ptr;      // start of each walk in walk_lin
  std::vector<int> walk_lin;      // voxel indices along each walk
  // row view (output-row order) and voxel-major transpose
  std::vector<i64> row_ptr;       // m+1
  std::vector<int> row_walk;      // walks of each row, in loop order
  std::vector<i64> col_ptr;       // n+1
  std::vector<int> col_ray;       // rays through each voxel, ascending ray
};

/* The Bresenham walk of one emitter-receiver pair, restated from the
 * recurrence at 3d/mat_mult.h:206-359: the major axis is the first of x, y, z
 * with the largest |d| (ties go to x, then y); start and end voxels are both
 * visited; the error terms start at 2*m_minor - L and a minor step happens
 * when the error is > 0. */
void bresenham(int x, int y, int z, int x1, int y1, int z1, int rvx, int rvy,
               std::vector<int> &out) {
  const int dx = x1 - x, dy = y1 - y, dz = z1 - z;
  const int xi = dx < 0 ? -1 : 1, mx = std::abs(dx);
  const int yi = dy < 0 ? -1 : 1, my = std::abs(dy);
  const int zi = dz < 0 ? -1 : 1, mz = std::abs(dz);
  auto emit = [&]() { out.push_back(x + y * rvx + z * rvx * rvy); };
  emit();
  if (mx >= my && mx >= mz) {
    int e1 = 2 * my - mx, e2 = 2 * mz - mx;
    for (int j = 1; j < mx + 1; ++j) {
      if (e1 > 0) { y += yi; e1 -= 2 * mx; }
      if (e2 > 0) { z += zi; e2 -= 2 * mx; }
      e1 += 2 * my; e2 += 2 * mz; x += xi;
      emit();
    }
  } else if (my >= mx && my >= mz) {
    int e1 = 2 * mx - my, e2 = 2 * mz - my;
    for (int j = 1; j < my + 1; ++j) {
      if (e1 > 0) { x += xi; e1 -= 2 * my; }
      if (e2 > 0) { z += zi; e2 -= 2 * my; }
      e1 += 2 * mx; e2 += 2 * mz; y += yi;
      emit();
    }
  } else {
    int e1 = 2 * mx - mz, e2 = 2 * my - mz;
    for (int j = 1; j < mz + 1; ++j) {
      if (e1 > 0) { x += xi; e1 -= 2 * mz; }
      if (e2 > 0) { y += yi; e2 -= 2 * mz; }
      e1 += 2 * mx; e2 += 2 * my; z += zi;
      emit();
    }
  }
}

void build_geometry(Op &op, const tv3_geometry &g) {
  op.kind = 2;
  op.m = (i64)g.num_mps * g.num_paths;
  op.n = (i64)g.rv_x * g.rv_y * g.rv_z;
  op.scale = g.scale_factor;
  op.walk_ptr.push_back(0);
  const int E = g.num_emitters, R = g.num_receivers;
  for (int i = 0; i < g.num_mps; ++i)
    for (int e = 0; e < E; ++e)
      for (int r = 0; r < R; ++r) {
        const int up = g.use_path[e * R + r];
        if (up == 0) continue;
        const int ray = i * g.num_paths + up - 1;
        bresenham(g.x_emitters[i * E + e], g.y_emitters[i * E + e],
                  g.z_emitters[i * E + e], g.x_receivers[i * R + r],
                  g.y_receivers[i * R + r], g.z_receivers[i * R + r], g.rv_x,
                  g.rv_y, op.walk_lin);
        op.walk_ray.push_back(ray);
        op.walk_ptr.push_back((i64)op.walk_lin.size());
      }
  op.nnz = (i64)op.walk_lin.size();
  const i64 nw = (i64)op.walk_ray.size();
  // rows
  op.row_ptr.assign(op.m + 1, 0);
  for (i64 w = 0; w < nw; ++w) op.row_ptr[op.walk_ray[w] + 1]++;
  for (i64 r = 0; r < op.m; ++r) op.row_ptr[r + 1] += op.row_ptr[r];
  op.row_walk.resize(nw);
  {
    std::vector<i64> fill(op.row_ptr.begin(), op.row_ptr.end() - 1);
    for (i64 w = 0; w < nw; ++w) op.row_walk[fill[op.walk_ray[w]]++] = (int)w;
  }
  // voxel-major transpose: rays in ascending output row, each walk in order
  op.col_ptr.assign(op.n + 1, 0);
  for (i64 k = 0; k < op.nnz; ++k) op.col_ptr[op.walk_lin[k] + 1]++;
  for (i64 v = 0; v < op.n; ++v) op.col_ptr[v + 1] += op.col_ptr[v];
  op.col_ray.resize(op.nnz);
  std::vector<i64> fill(op.col_ptr.begin(), op.col_ptr.end() - 1);
  for (i64 r = 0; r < op.m; ++r)
    for (i64 t = op.row_ptr[r]; t < op.row_ptr[r + 1]; ++t) {
      const int w = op.row_walk[t];
      for (i64 k = op.walk_ptr[w]; k < op.walk_ptr[w + 1]; ++k)
        op.col_ray[fill[op.walk_lin[k]]++] = (int)r;
    }
}

/* get_g_mult for a dense A (2d/tval3_cpu_2d.cpp:361-394, single layer;
 * 3d/tval3_cpu_3d.cpp:498-529): row blocks of r = max(submat_mb MiB /
 * sizeof(T) / n, 1) rows; per block Au_blk = A_blk U and g_tmp = A_blk^T
 * Au_blk (each a gemv: FP64 sums rounded to T), then g += g_tmp (cblas_axpy
 * with alpha 1: one rounding per block, in block order). */
template <class T>
void dense_AtA_blocked(const Op &op, const T *U, T *Au, T *g, int submat_mb) {
  const i64 m = op.m, n = op.n;
  const i64 elems = (i64)submat_mb * 1024 * 1024 / (i64)sizeof(T);
  const i64 r = std::max<i64>(elems / n, 1);
  const float *A = op.A;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < m; ++i) {
    double s = 0.0;
    const float *row = A + i * n;
    for (i64 j = 0; j < n; ++j) s += (double)row[j] * (double)U[j];
    Au[i] = (T)s;
  }
  const i64 cb = 256;
#pragma omp parallel for schedule(static)
  for (i64 j0 = 0; j0 < n; j0 += cb) {
    const i64 j1 = std::min(n, j0 + cb);
    T gacc[256];
    double acc[256];
    for (i64 j = j0; j < j1; ++j) gacc[j - j0] = (T)0;
    for (i64 y0 = 0; y0 < m; y0 += r) {
      const i64 y1 = std::min(m, y0 + r);
      for (i64 j = j0; j < j1; ++j) acc[j - j0] = 0.0;
      for (i64 i = y0; i < y1; ++i) {
        const double ai = (double)Au[i];
        const float *row = A + i * n;
        for (i64 j = j0; j < j1; ++j) acc[j - j0] += (double)row[j] * ai;
      }
      for (i64 j = j0; j < j1; ++j) gacc[j - j0] = gacc[j - j0] + (T)acc[j - j0];
    }
    for (i64 j = j0; j < j1; ++j) g[j] = gacc[j - j0];
  }
}

/* y = A x or A^T x, overwriting y (gemv with beta = 0).  T is the vector
 * type; geometry products are formed in it (mult_elementN/T, 3d/mat_mult.h:
 * 121-139), all sums in FP64. */
template <class T>
void gemv(const Op &op, bool trans, const T *x, T *y) {
  const i64 m = op.m, n = op.n;
  if (op.kind == 0) {
    const float *A = op.A;
    if (!trans) {
#pragma omp parallel for schedule(static)
      for (i64 i = 0; i < m; ++i) {
        double s = 0.0;
        const float *row = A + i * n;
        for (i64 j = 0; j < n; ++j) s += (double)row[j] * (double)x[j];
        y[i] = (T)s;
      }
    } else {
      const i64 cb = 256;
#pragma omp parallel for schedule(static)
      for (i64 j0 = 0; j0 < n; j0 += cb) {
        const i64 j1 = std::min(n, j0 + cb);
        double acc[256];
        for (i64 j = j0; j < j1; ++j) acc[j - j0] = 0.0;
        for (i64 i = 0; i < m; ++i) {
          const double xi = (double)x[i];
          const float *row = A + i * n;
          for (i64 j = j0; j < j1; ++j) acc[j - j0] += (double)row[j] * xi;
        }
        for (i64 j = j0; j < j1; ++j) y[j] = (T)acc[j - j0];
      }
    }
  } else if (op.kind == 1) {
    const int *ptr = trans ? op.csc_ptr : op.csr_ptr;
    const int *ind = trans ? op.csc_ind : op.csr_ind;
    const float *val = trans ? op.csc_val : op.csr_val;
    const i64 rows = trans ? n : m;
#pragma omp parallel for schedule(dynamic, 64)
    for (i64 r = 0; r < rows; ++r) {
      double s = 0.0;
      for (int k = ptr[r]; k < ptr[r + 1]; ++k)
        s += (double)val[k] * (double)x[ind[k]];
      y[r] = (T)s;
    }
  } else {
    const T sc = (T)op.scale;
    if (!trans) {
#pragma omp parallel for schedule(dynamic, 64)
      for (i64 r = 0; r < m; ++r) {
        double s = 0.0;
        for (i64 t = op.row_ptr[r]; t < op.row_ptr[r + 1]; ++t) {
          const int w = op.row_walk[t];
          for (i64 k = op.walk_ptr[w]; k < op.walk_ptr[w + 1]; ++k)
            s += (double)(x[op.walk_lin[k]] * sc);
        }
        y[r] = (T)s;
      }
    } else {
#pragma omp parallel for schedule(dynamic, 256)
      for (i64 v = 0; v < n; ++v) {
        double s = 0.0;
        for (i64 k = op.col_ptr[v]; k < op.col_ptr[v + 1]; ++k)
          s += (double)(x[op.col_ray[k]] * sc);
        y[v] = (T)s;
      }
    }
  }
}

/* Complex products (the complex instantiations of gemv, 2d/mat_mult.h:15-91
 * with complex A, and of get_g_mult's blocked product): y = A x takes the
 * terms a x, y = A^H x (CblasConjTrans, 2d:308, 392, 774) the terms conj(a)
 * x, each component of a term one FP64 rounding of two FP64 products, all
 * sums in FP64 and rounded once per component.  The geometry projector has
 * real weights: each component is multiplied by the scale factor in the
 * component type. */
struct CAcc {
  double re = 0.0, im = 0.0;
};
inline void cadd(CAcc &s, double ar, double ai, double xr, double xi, bool conj) {
  if (!conj) {
    s.re += ar * xr - ai * xi;
    s.im += ar * xi + ai * xr;
  } else {
    s.re += ar * xr + ai * xi;
    s.im += ar * xi - ai * xr;
  }
}

template <class R>
void dense_AtA_blocked(const Op &op, const std::complex<R> *U, std::complex<R> *Au,
                       std::complex<R> *g, int submat_mb) {
  typedef std::complex<R> T;
  const i64 m = op.m, n = op.n;
  const i64 elems = (i64)submat_mb * 1024 * 1024 / (i64)sizeof(T);
  const i64 r = std::max<i64>(elems / n, 1);
  const double *A = op.Az;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < m; ++i) {
    CAcc s;
    const double *row = A + 2 * i * n;
    for (i64 j = 0; j < n; ++j)
      cadd(s, row[2 * j], row[2 * j + 1], (double)U[j].real(), (double)U[j].imag(), false);
    Au[i] = T((R)s.re, (R)s.im);
  }
  const i64 cb = 256;
#pragma omp parallel for schedule(static)
  for (i64 j0 = 0; j0 < n; j0 += cb) {
    const i64 j1 = std::min(n, j0 + cb);
    T gacc[256];
    CAcc acc[256];
    for (i64 j = j0; j < j1; ++j) gacc[j - j0] = T(0, 0);
    for (i64 y0 = 0; y0 < m; y0 += r) {
      const i64 y1 = std::min(m, y0 + r);
      for (i64 j = j0; j < j1; ++j) acc[j - j0] = CAcc();
      for (i64 i = y0; i < y1; ++i) {
        const double xr = (double)Au[i].real(), xi = (double)Au[i].imag();
        const double *row = A + 2 * i * n;
        for (i64 j = j0; j < j1; ++j) cadd(acc[j - j0], row[2 * j], row[2 * j + 1], xr, xi, true);
      }
      for (i64 j = j0; j < j1; ++j)
        gacc[j - j0] = T(gacc[j - j0].real() + (R)acc[j - j0].re, gacc[j - j0].imag() + (R)acc[j - j0].im);
    }
    for (i64 j = j0; j < j1; ++j) g[j] = gacc[j - j0];
  }
}

template <class R>
void gemv(const Op &op, bool trans, const std::complex<R> *x, std::complex<R> *y) {
  typedef std::complex<R> T;
  const i64 m = op.m, n = op.n;
  if (op.kind == 0) {
    const double *A = op.Az;
    if (!trans) {
#pragma omp parallel for schedule(static)
      for (i64 i = 0; i < m; ++i) {
        CAcc s;
        const double *row = A + 2 * i * n;
        for (i64 j = 0; j < n; ++j)
          cadd(s, row[2 * j], row[2 * j + 1], (double)x[j].real(), (double)x[j].imag(), false);
        y[i] = T((R)s.re, (R)s.im);
      }
    } else {
      const i64 cb = 256;
#pragma omp parallel for schedule(static)
      for (i64 j0 = 0; j0 < n; j0 += cb) {
        const i64 j1 = std::min(n, j0 + cb);
        CAcc acc[256];
        for (i64 i = 0; i < m; ++i) {
          const double xr = (double)x[i].real(), xi = (double)x[i].imag();
          const double *row = A + 2 * i * n;
          for (i64 j = j0; j < j1; ++j) cadd(acc[j - j0], row[2 * j], row[2 * j + 1], xr, xi, true);
        }
        for (i64 j = j0; j < j1; ++j) y[j] = T((R)acc[j - j0].re, (R)acc[j - j0].im);
      }
    }
  } else if (op.kind == 1) {
    const int *ptr = trans ? op.csc_ptr : op.csr_ptr;
    const int *ind = trans ? op.csc_ind : op.csr_ind;
    const double *val = trans ? op.csc_valz : op.csr_valz;
    const i64 rows = trans ? n : m;
#pragma omp parallel for schedule(dynamic, 64)
    for (i64 r = 0; r < rows; ++r) {
      CAcc s;
      for (int k = ptr[r]; k < ptr[r + 1]; ++k)
        cadd(s, val[2 * k], val[2 * k + 1], (double)x[ind[k]].real(), (double)x[ind[k]].imag(), trans);
      y[r] = T((R)s.re, (R)s.im);
    }
  } else {
    const R sc = (R)op.scale;
    if (!trans) {
#pragma omp parallel for schedule(dynamic, 64)
      for (i64 r = 0; r < m; ++r) {
        double sr = 0.0, si = 0.0;
        for (i64 t = op.row_ptr[r]; t < op.row_ptr[r + 1]; ++t) {
          const int w = op.row_walk[t];
          for (i64 k = op.walk_ptr[w]; k < op.walk_ptr[w + 1]; ++k) {
            const T v = x[op.walk_lin[k]];
            sr += (double)(v.real() * sc);
            si += (double)(v.imag() * sc);
          }
        }
        y[r] = T((R)sr, (R)si);
      }
    } else {
#pragma omp parallel for schedule(dynamic, 256)
      for (i64 v = 0; v < n; ++v) {
        double sr = 0.0, si = 0.0;
        for (i64 k = op.col_ptr[v]; k < op.col_ptr[v + 1]; ++k) {
          const T u = x[op.col_ray[k]];
          sr += (double)(u.real() * sc);
          si += (double)(u.imag() * sc);
        }
        y[v] = T((R)sr, (R)si);
      }
    }
  }
}

/* ---------------------------------------------------------------------- */
/* finite differences (x is the contiguous axis; periodic boundaries)      */
/* ---------------------------------------------------------------------- */
struct Dims {
  int q, p, r;  // dim_x, dim_y, dim_z
  i64 n() const { return (i64)q * p * r; }
};

/* D: 2d/tval3_cpu_2d.cpp:268-278, 3d/tval3_cpu_3d.cpp:365-389 */
template <class T>
void Dop(const Dims &d, const T *U, T *Ux, T *Uy, T *Uz) {
  const i64 q = d.q, p = d.p, r = d.r, s = q * p, n = d.n();
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) {
    const i64 x = i % q, y = (i / q) % p, z = i / s;
    Ux[i] = (x < q - 1 ? U[i + 1] : U[i - (q - 1)]) - U[i];
    Uy[i] = (y < p - 1 ? U[i + q] : U[i - (p - 1) * q]) - U[i];
    if (r > 1) Uz[i] = (z < r - 1 ? U[i + s] : U[i - (r - 1) * s]) - U[i];
  }
}

/* Dt: 2d/tval3_cpu_2d.cpp:280-293, 3d/tval3_cpu_3d.cpp:394-425.  The
 * reference builds res with a sequence of vector passes; each entry sees the
 * same additions in the same order here.  For dim_z >= 2 the reference's bulk
 * y-pass runs over all i >= dim_x, across slice boundaries, and the per-slice
 * wrap pass then also applies to the first row of every slice: the first row
 * of slice z >= 1 receives an extra +Y[i - dim_x] - Y[i].  exact = 1 omits
 * that extra term. */
template <class T>
void Dtop(const Dims &d, bool exact, const T *X, const T *Y, const T *Z,
          T *res) {
  const i64 q = d.q, p = d.p, r = d.r, s = q * p, n = d.n();
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) {
    const i64 x = i % q, y = (i / q) % p;
    T t = (x > 0 ? X[i - 1] : X[i + q - 1]) - X[i];
    if (r == 1) {
      if (y > 0) {
        t = t + Y[i - q];
        t = t - Y[i];
      } else {
        t = t + Y[i + (p - 1) * q];
        t = t - Y[i];
      }
    } else {
      if (i >= q && (y > 0 || !exact)) {  // bulk pass, i >= dim_x
        t = t + Y[i - q];
        t = t - Y[i];
      }
      if (y == 0) {  // per-slice wrap pass
        t = t + Y[i + (p - 1) * q];
        t = t - Y[i];
      }
      if (g_plant == 5 && exact) {
      } else if (i >= s) {
        t = t + Z[i - s];
        t = t - Z[i];
      } else {
        t = t + Z[i + (r - 1) * s];
        t = t - Z[i];
      }
    }
    res[i] = t;
  }
}

/* shrinkage: 2d/tval3_cpu_2d.cpp:588-611 (anisotropic, the reference's only
 * real variant); iso = 1 is the isotropic group soft-threshold of the TVAL3
 * method, which the reference does not implement.  Returns sum in FP64. */
template <class T, class TS>
double shrink(const Dims &d, TS beta, bool iso, const T *Ux, const T *Uy,
              const T *Uz, const T *sx, const T *sy, const T *sz, T *Wx,
              T *Wy, T *Wz) {
  const i64 n = d.n();
  const bool z3 = d.r > 1;
  const TS ib = (TS)1 / beta;
  return chunked_sum(n, [&](i64 i) -> double {
    const T ubx = (T)((TS)Ux[i] - (TS)sx[i] / beta);
    const T uby = (T)((TS)Uy[i] - (TS)sy[i] / beta);
    const T ubz = z3 ? (T)((TS)Uz[i] - (TS)sz[i] / beta) : (T)0;
    if (!iso) {
      TS wx = std::abs((TS)ubx) - ib, wy = std::abs((TS)uby) - ib;
      wx = (wx >= 0) ? wx : (TS)0;
      wy = (wy >= 0) ? wy : (TS)0;
      Wx[i] = (T)((ubx >= 0) ? wx : -wx);
      Wy[i] = (T)((uby >= 0) ? wy : -wy);
      TS t = wx + wy;
      if (z3) {
        TS wz = std::abs((TS)ubz) - ib;
        wz = (wz >= 0) ? wz : (TS)0;
        Wz[i] = (T)((ubz >= 0) ? wz : -wz);
        if (g_plant != 12) t = t + wz;
      }
      return (double)t;
    }
    TS ss = (TS)ubx * (TS)ubx + (TS)uby * (TS)uby;
    if (z3) ss = ss + (TS)ubz * (TS)ubz;
    const TS rr = std::sqrt(ss);
    TS w = rr - ib;
    w = (w >= 0) ? w : (TS)0;
    Wx[i] = (rr > 0) ? (T)(w * (TS)ubx / rr) : (T)0;
    Wy[i] = (rr > 0) ? (T)(w * (TS)uby / rr) : (T)0;
    if (z3) Wz[i] = (rr > 0) ? (T)(w * (TS)ubz / rr) : (T)0;
    return (double)w;
  });
}

/* complex shrinkage: 2d/tval3_cpu_2d.cpp:614-649, 3d/tval3_cpu_3d.cpp:876-947.
 * Per direction: ubar = u - sigma/beta per component (rounded to R),
 * |ubar| = sqrt(re*re + im*im) in R, w = max(|ubar| - 1/beta, 0) in R and
 * W = w * ubar / |ubar| per component in R (0 where |ubar| = 0).  The
 * reference's 3-D version takes Wz's components from the y direction's ubar
 * (3d:934-937); wz_fix = 1 takes the z direction's.  Returns the FP64 sum of
 * the per-voxel sums of w (each in R). */
template <class R, class TS>
double shrink_c(const Dims &d, TS beta, bool wz_fix, const std::complex<R> *Ux,
                const std::complex<R> *Uy, const std::complex<R> *Uz,
                const std::complex<R> *sx, const std::complex<R> *sy,
                const std::complex<R> *sz, std::complex<R> *Wx,
                std::complex<R> *Wy, std::complex<R> *Wz) {
  typedef std::complex<R> T;
  const i64 n = d.n();
  const bool z3 = d.r > 1;
  const TS ib = (TS)1 / beta;
  auto bar = [&](const T &u, const T &sg) {
    return T((R)((TS)u.real() - (TS)sg.real() / beta), (R)((TS)u.imag() - (TS)sg.imag() / beta));
  };
  auto mag = [](const T &b) -> R { return std::sqrt(b.real() * b.real() + b.imag() * b.imag()); };
  auto thr = [&](R a) -> R {
    R t = (R)((TS)a - ib);
    return (t >= 0) ? t : (R)0;
  };
  auto dirW = [](R w, const T &b, R a) {
    return (a > 0) ? T(w * b.real() / a, w * b.imag() / a) : T(0, 0);
  };
  return chunked_sum(n, [&](i64 i) -> double {
    const T bx = bar(Ux[i], sx[i]), by = bar(Uy[i], sy[i]);
    const R ax = mag(bx), ay = mag(by);
    const R wx = thr(ax), wy = thr(ay);
    Wx[i] = dirW(wx, bx, ax);
    Wy[i] = dirW(wy, by, ay);
    R t = wx + wy;
    if (z3) {
      const T bz = bar(Uz[i], sz[i]);
      const R az = mag(bz), wz = thr(az);
      Wz[i] = dirW(wz, wz_fix ? bz : by, az);
      t = t + wz;
    }
    return (double)t;
  });
}

/* ---------------------------------------------------------------------- */
/* the solver                                                              */
/* ---------------------------------------------------------------------- */
template <class TC, class TS>
struct Solver {
  typedef std::vector<TC> V;
  // component type: f, C, lam4, lam5 are T_comp in the reference; with
  // complex T_comp they are real-valued (imaginary parts stay 0), so they
  // are kept as their real component
  typedef typename VT<TC>::R TR;
  static const bool kCplx = VT<TC>::cplx;
  const Op &A;
  const tv3_options &o;
  Dims dm;
  i64 n, m;
  bool z3;
  // 2-D multi-layer mode (2d/tval3_cpu_2d.cpp:889-1016 pass layers =
  // b.len / A.dim_y): U holds `layers` images of A.n pixels one after the
  // other and b `layers` measurement vectors of A.m; every product applies A
  // to each layer (2d/mat_mult.h:22-35, 68-90)
  int layers;
  // options in T_scal
  TS o_c, o_gamma, o_rate_gam, o_rate_cnt, o_gam, o_tol, o_tol_inn;
  // penalty parameters, multipliers, Lagrangian terms (tval3_data)
  TS mu, beta, muDbeta;
  V delta, sx, sy, sz;
  TR f;
  TS lam1, lam2, lam3;
  TR lam4, lam5;
  V Up, dU, uup, Ux, Uxp, dUx, Uy, Uyp, dUy, Uz, Uzp, dUz, Wx, Wy, Wz, Atb,
      Au, Aup, dAu, d, g, gp, dg, g2, g2p, dg2;
  unsigned currInitIteration;
  TS currSkipValue;
  V last_Au;
  // per-iteration trace fields
  int bt_cnt, bt_fallback;
  TS last_tau, last_alpha;

  Solver(const Op &A_, const tv3_options &o_, Dims d_)
      : A(A_), o(o_), dm(d_), n(d_.n()), m(A_.m * (d_.n() / A_.n)), z3(d_.r > 1),
        layers((int)(d_.n() / A_.n)) {
    o_c = (TS)o.c; o_gamma = (TS)o.gamma; o_rate_gam = (TS)o.rate_gam;
    o_rate_cnt = (TS)o.rate_cnt; o_gam = (TS)o.gam; o_tol = (TS)o.tol;
    o_tol_inn = (TS)o.tol_inn;
    mu = (TS)o.mu0; beta = (TS)o.beta0; muDbeta = 0;
    f = 0; lam1 = lam2 = lam3 = 0; lam4 = lam5 = 0;
    const i64 nz = z3 ? n : 0;
    delta.assign(m, (TC)0); last_Au.assign(m, (TC)0);
    Au.assign(m, (TC)0); Aup.assign(m, (TC)0); dAu.assign(m, (TC)0);
    for (V *v : {&sx, &sy, &Up, &dU, &uup, &Ux, &Uxp, &dUx, &Uy, &Uyp, &dUy,
                 &Wx, &Wy, &Atb, &d, &g, &gp, &dg, &g2, &g2p, &dg2})
      v->assign(n, (TC)0);
    for (V *v : {&sz, &Uz, &Uzp, &dUz, &Wz}) v->assign(nz, (TC)0);
    currInitIteration = 0; currSkipValue = 0;
    bt_cnt = bt_fallback = 0; last_tau = last_alpha = 0;
  }

  TC *P(V &v) { return v.empty() ? nullptr : v.data(); }

  void D_(const V &U, V &X, V &Y, V &Z) { Dop(dm, U.data(), P(X), P(Y), P(Z)); }
  void Dt_(V &res, const V &X, const V &Y, const V &Z) {
    Dtop(dm, o.dt_exact != 0, X.data(), Y.data(), z3 ? Z.data() : nullptr,
         res.data());
  }

  /* scaleb: 2d/tval3_cpu_2d.cpp:228-266 */
  TS scaleb(V &b) {
    TS scl = 1;
    if (m > 0) {
      const TS th1 = (TS)0.5, th2 = (TS)1.5;
      TC bmin = b[0], bmax = bmin;
      TS bmin_abs = modulus(bmin), bmax_abs = bmin_abs;
      for (i64 i = 0; i < m; ++i) {
        const TC val = b[i];
        const TS va = modulus(val);
        if (va < bmin_abs) { bmin = val; bmin_abs = va; }
        if (va > bmax_abs) { bmax = val; bmax_abs = va; }
      }
      const TS b_dif = modulus(bmax - bmin);
      if (b_dif < th1) {
        scl = th1 / b_dif;
        scal(m, (TR)scl, b.data());
      } else if (b_dif > th2) {
        scl = th2 / b_dif;
        scal(m, (TR)scl, b.data());
      }
    }
    return scl;
  }

  /* lam2 and lam4 from V = U_dir - W_dir (get_g / update_g / update_W) */
  void lam24(V &Vx, V &Vy, V &Vz) {
    vec_sub(n, Ux.data(), Wx.data(), Vx.data());
    vec_sub(n, Uy.data(), Wy.data(), Vy.data());
    if (z3) vec_sub(n, Uz.data(), Wz.data(), Vz.data());
    lam2 = (TS)sdot(n, Vx.data(), Vx.data());
    lam2 += (TS)sdot(n, Vy.data(), Vy.data());
    if (z3 && g_plant != 2) lam2 += (TS)sdot(n, Vz.data(), Vz.data());
    lam4 = sdot(n, sx.data(), Vx.data());
    lam4 += sdot(n, sy.data(), Vy.data());
    if (z3 && g_plant != 3) lam4 += sdot(n, sz.data(), Vz.data());
  }
  void lam35(const V &b) {
    V Aub(m);
    vec_sub(m, Au.data(), b.data(), Aub.data());
    lam3 = (TS)sdot(m, Aub.data(), Aub.data());
    lam5 = sdot(m, delta.data(), Aub.data());
  }
  TR f_full() {
    return (TR)((((lam1 + beta / 2 * lam2) + mu / 2 * lam3) - (TS)lam4) -
                   (TS)lam5);
  }

  /* gemv with layers: A applied to each layer (the reference's gemm /
   * cscmm computes the same products; each output is an FP64 sum rounded
   * once, as for one layer) */
  void gemvL(bool trans, const TC *x, TC *y) {
    ++(trans ? g_nbwd : g_nfwd);
    const i64 in = trans ? A.m : A.n, out = trans ? A.n : A.m;
    for (int l = 0; l < layers; ++l) gemv(A, trans, x + l * in, y + l * out);
  }

  /* get_g_mult: 2d/tval3_cpu_2d.cpp:298-357 (skipping state machine) and
   * 359-394 (dense: always both products; blocked for one layer, two plain
   * products for several) */
  void get_g_mult(const V &U, bool force_mul) {
    if (A.kind == 0) {
      if (layers == 1) {
        ++g_nfwd;
        ++g_nbwd;
        dense_AtA_blocked(A, U.data(), Au.data(), g.data(), o.submat_mb > 0 ? o.submat_mb : 5);
      } else {
        gemvL(false, U.data(), Au.data());
        gemvL(true, Au.data(), g.data());
      }
      return;
    }
    gemvL(false, U.data(), Au.data());
    if (o.skipMulRatio == 0) {
      gemvL(true, Au.data(), g.data());
      return;
    }
    V deltaAu(Au);
    axpy(m, (TR)-1, last_Au.data(), deltaAu.data());
    const TS nrm_deltaAu = snrm2(m, deltaAu.data());
    const TS nrm_Au = snrm2(m, Au.data());
    const TS relChange = nrm_deltaAu / nrm_Au;
    const TS ratio = (TS)o.skipMulRatio, maxrel = (TS)o.maxRelativeChange;
    if (force_mul) {
      last_Au = Au; gemvL(true, Au.data(), g.data());
      if (g_plant != 11) currSkipValue = 0;
    } else if (currInitIteration < o.numInitIterations) {
      last_Au = Au; gemvL(true, Au.data(), g.data());
      currInitIteration++; currSkipValue = 0;
    } else if (currSkipValue < 1) {
      last_Au = Au; gemvL(true, Au.data(), g.data());
      if (g_plant != 10) currSkipValue += ratio;
    } else if (relChange > maxrel) {
      last_Au = Au; gemvL(true, Au.data(), g.data());
    } else {
      currSkipValue -= 1;
    }
  }

  /* get_g: 2d/tval3_cpu_2d.cpp:396-423 */
  void get_g(const V &U, const V &b, bool force_mul) {
    get_g_mult(U, force_mul);
    axpy(n, (TR)-1, Atb.data(), g.data());
    V Vx(n), Vy(n), Vz(z3 ? n : 0);
    vec_sub(n, Ux.data(), Wx.data(), Vx.data());
    vec_sub(n, Uy.data(), Wy.data(), Vy.data());
    if (z3) vec_sub(n, Uz.data(), Wz.data(), Vz.data());
    Dt_(g2, Vx, Vy, Vz);
    lam24(Vx, Vy, Vz);
    lam35(b);
    f = f_full();
  }

  /* get_tau: 2d/tval3_cpu_2d.cpp:425-453 */
  TS get_tau(bool fst_iter) {
    TS tau;
    if (fst_iter) {
      V dx(n), dy(n), dz(z3 ? n : 0);
      D_(d, dx, dy, dz);
      TR dDd = sdot(n, dx.data(), dx.data()) + sdot(n, dy.data(), dy.data());
      if (z3 && g_plant != 1) dDd = dDd + sdot(n, dz.data(), dz.data());
      const TR dd = sdot(n, d.data(), d.data());
      V Ad(m);
      gemvL(false, d.data(), Ad.data());
      const TR Add = g_plant == 14 ? (TR)0 : sdot(m, Ad.data(), Ad.data());
      tau = std::abs((TS)dd / ((TS)dDd + muDbeta * (TS)Add));
    } else {
      vec_sub(n, g.data(), gp.data(), dg.data());
      vec_sub(n, g2.data(), g2p.data(), dg2.data());
      const TR ss = sdot(n, uup.data(), uup.data());
      V tmp(dg2);
      if (g_plant != 6) axpy(n, (TR)muDbeta, dg.data(), tmp.data());
      const TR sy_ = sdot(n, uup.data(), tmp.data());
      if constexpr (kCplx) {
        // abs(ss / sy) of complex values (ss real-valued): |ss| / |sy| in
        // FP64, rounded to the component type
        const TR sy_i = sdot_im(n, uup.data(), tmp.data());
        const double den = modulus(std::complex<double>((double)sy_, (double)sy_i));
        tau = (TS)(TR)(std::abs((double)ss) / den);
      } else {
        tau = (TS)std::abs(ss / sy_);
      }
    }
    return tau;
  }

  /* nonneg: 2d/tval3_cpu_2d.cpp:455-479 (complex: the real part clamped,
   * the imaginary part set to 0) */
  static void clamp0(float &u) { u = (u < 0) ? 0.0f : u; }
  static void clamp0(double &u) { u = (u < 0) ? 0.0 : u; }
  template <class R>
  static void clamp0(std::complex<R> &u) { u = std::complex<R>(u.real() < 0 ? (R)0 : u.real(), (R)0); }
  void nonneg(V &U) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) clamp0(U[i]);
  }
  TC finish(float u, TS scl) const { return (TC)((TS)u / scl); }
  TC finish(double u, TS scl) const { return (TC)((TS)u / scl); }
  template <class R>
  TC finish(const std::complex<R> &u, TS scl) const {
    const R im = o.isreal ? (R)0 : u.imag();
    return TC((R)((TS)u.real() / scl), (R)((TS)im / scl));
  }

  /* descend: 2d/tval3_cpu_2d.cpp:481-493 */
  void descend(V &U, const V &b, TS tau, bool force_mul) {
    axpy(n, (TR)(-1 * tau), d.data(), U.data());
    if (o.nonneg) nonneg(U);
    D_(U, Ux, Uy, Uz);
    get_g(U, b, force_mul);
  }

  /* update_g: 2d/tval3_cpu_2d.cpp:495-524 */
  void update_g(TS alpha, V &U, const V &b) {
    vec_add<TC, TS>(n, gp.data(), alpha, dg.data(), g.data());
    vec_add<TC, TS>(n, g2p.data(), alpha, dg2.data(), g2.data());
    vec_add<TC, TS>(n, Up.data(), alpha, dU.data(), U.data());
    vec_add<TC, TS>(m, Aup.data(), alpha, dAu.data(), Au.data());
    vec_add<TC, TS>(n, Uxp.data(), alpha, dUx.data(), Ux.data());
    vec_add<TC, TS>(n, Uyp.data(), alpha, dUy.data(), Uy.data());
    if (z3) vec_add<TC, TS>(n, Uzp.data(), alpha, dUz.data(), Uz.data());
    V Vx(n), Vy(n), Vz(z3 ? n : 0);
    lam24(Vx, Vy, Vz);
    lam35(b);
    f = f_full();
  }

  /* min_u: 2d/tval3_cpu_2d.cpp:526-576 */
  void min_u(V &U, TS &gam, const V &b, TR C, bool fst_iter) {
    TS tau = get_tau(fst_iter);
    Up = U; gp = g; g2p = g2; Aup = Au; Uxp = Ux; Uyp = Uy;
    if (z3) Uzp = Uz;
    descend(U, b, tau, false);
    TS alpha = 1;
    vec_sub(n, U.data(), Up.data(), dU.data());
    const TS c_armij = g_plant == 7 ? (TS)0 : (TS)sdot(n, d.data(), d.data()) * tau * o_c * beta;
    bt_cnt = 0; bt_fallback = 0;
    if ((TS)std::abs(f) > std::abs((TS)C - alpha * c_armij)) {
      vec_sub(n, g.data(), gp.data(), dg.data());
      vec_sub(n, g2.data(), g2p.data(), dg2.data());
      vec_sub(m, Au.data(), Aup.data(), dAu.data());
      vec_sub(n, Ux.data(), Uxp.data(), dUx.data());
      vec_sub(n, Uy.data(), Uyp.data(), dUy.data());
      if (z3) vec_sub(n, Uz.data(), Uzp.data(), dUz.data());
      int cnt = 0;
      while ((TS)std::abs(f) > std::abs((TS)C - alpha * c_armij)) {
        if (cnt == 5) {
          if (g_plant != 8) gam *= o_rate_gam;
          tau = get_tau(true);
          if (g_plant != 9) U = Up;
          descend(U, b, tau, true);
          bt_fallback = 1;
          break;
        }
        alpha *= o_gamma;
        update_g(alpha, U, b);
        cnt++;
      }
      bt_cnt = cnt;
    }
    last_tau = tau;
    last_alpha = alpha;
  }

  /* get_gradient: 2d/tval3_cpu_2d.cpp:578-585 */
  void get_gradient(const V &DtsAtd) {
    d = DtsAtd;
    axpy(n, (TR)1, g2.data(), d.data());
    axpy(n, (TR)muDbeta, g.data(), d.data());
  }

  void shrinkage() { shrinkage_(Ux.data()); }
  template <class T>
  void shrinkage_(const T *) {
    lam1 = (TS)shrink<TC, TS>(dm, beta, o.tv_iso != 0, Ux.data(), Uy.data(),
                          P(Uz), sx.data(), sy.data(), P(sz), Wx.data(),
                          Wy.data(), P(Wz));
  }
  template <class R>
  void shrinkage_(const std::complex<R> *) {
    // the complex sum is T_mat (2d:617): rounded to the component type
    lam1 = (TS)(TR)shrink_c<R, TS>(dm, beta, o.dt_exact != 0, Ux.data(), Uy.data(), P(Uz),
                                   sx.data(), sy.data(), P(sz), Wx.data(), Wy.data(), P(Wz));
  }

  /* update_W: 2d/tval3_cpu_2d.cpp:651-671 */
  void update_W() {
    f = (TR)((TS)f - ((lam1 + beta / 2 * lam2) - (TS)lam4));
    shrinkage();
    V Vx(n), Vy(n), Vz(z3 ? n : 0);
    vec_sub(n, Ux.data(), Wx.data(), Vx.data());
    vec_sub(n, Uy.data(), Wy.data(), Vy.data());
    if (z3) vec_sub(n, Uz.data(), Wz.data(), Vz.data());
    Dt_(g2, Vx, Vy, Vz);
    lam24(Vx, Vy, Vz);
    f = (TR)((TS)f + ((lam1 + beta / 2 * lam2) - (TS)lam4));
  }

  /* update_mlp: 2d/tval3_cpu_2d.cpp:673-694 */
  void update_mlp(const V &b) {
    f = f + (lam4 + lam5);
    V Vx(n), Vy(n), Vz(z3 ? n : 0);
    vec_sub(n, Ux.data(), Wx.data(), Vx.data());
    vec_sub(n, Uy.data(), Wy.data(), Vy.data());
    if (z3) vec_sub(n, Uz.data(), Wz.data(), Vz.data());
    const TR nb = (TR)(-1 * beta);
    axpy(n, nb, Vx.data(), sx.data());
    axpy(n, nb, Vy.data(), sy.data());
    if (z3 && g_plant != 4) axpy(n, nb, Vz.data(), sz.data());
    lam4 = sdot(n, sx.data(), Vx.data());
    lam4 += sdot(n, sy.data(), Vy.data());
    if (z3) lam4 += sdot(n, sz.data(), Vz.data());
    V Aub(m);
    vec_sub(m, Au.data(), b.data(), Aub.data());
    axpy(m, (TR)(g_plant == 13 ? mu : -1 * mu), Aub.data(), delta.data());
    lam5 = sdot(m, delta.data(), Aub.data());
    f = f - (lam4 + lam5);
  }

  /* driver: 2d/tval3_cpu_2d.cpp:752-886 */
  void run(const TC *b_in, TC *U_out, const TC *Ut, tv3_info *info,
           tv3_trace_rec *trace, int trace_cap) {
    V bs(b_in, b_in + m);
    const TS muf = (TS)o.mu, betaf = (TS)o.beta;
    TS beta0 = 0;
    V DtsAtd(n, (TC)0);
    TS gam = o_gam;
    const TS scl = scaleb(bs);
    gemvL(true, bs.data(), Atb.data());
    V U(Atb);
    if (mu > muf) mu = muf;
    if (beta > betaf) beta = betaf;
    muDbeta = mu / beta;
    V rcdU(U), UrcdU(n);
    D_(U, Ux, Uy, Uz);
    shrinkage();
    get_g(U, bs, true);
    get_gradient(DtsAtd);
    TS Q = 1, Qp;
    TR C = f;
    int count_outer = 0, count_total = 0;
    bool fst_iter;
    TS RelChg, RelChgOut = 0, nrmup = 0;
    while ((count_outer < o.maxcnt) && (count_total < o.maxit)) {
      fst_iter = true;
      while (count_total < o.maxit) {
        min_u(U, gam, bs, C, fst_iter);
        update_W();
        Qp = Q, Q = gam * Qp + 1;
        C = (TR)((gam * Qp * (TS)C + (TS)f) / Q);
        vec_sub(n, U.data(), Up.data(), uup.data());
        get_gradient(DtsAtd);
        count_total++;
        nrmup = (TS)snrm2(n, Up.data());
        RelChg = (TS)snrm2(n, uup.data()) / nrmup;
        if (trace && count_total <= trace_cap) {
          tv3_trace_rec &t = trace[count_total - 1];
          t.iter = count_total; t.outer = count_outer;
          t.backtracks = bt_cnt; t.fallback = bt_fallback;
          t.tau = last_tau; t.alpha = last_alpha; t.f = f; t.C = C;
          t.rel_chg = RelChg; t.lam1 = lam1; t.lam2 = lam2; t.lam3 = lam3;
          t.lam4 = lam4; t.lam5 = lam5; t.mu = mu; t.beta = beta;
        }
        if ((RelChg < o_tol_inn) && (!fst_iter)) break;
        fst_iter = false;
      }
      count_outer++;
      vec_sub(n, U.data(), rcdU.data(), UrcdU.data());
      RelChgOut = (TS)snrm2(n, UrcdU.data()) / nrmup;
      rcdU = U;
      if (RelChgOut < o_tol) break;
      update_mlp(bs);
      beta0 = beta;
      beta *= o_rate_cnt;
      mu *= o_rate_cnt;
      if (beta > betaf) beta = betaf;
      if (mu > muf) mu = muf;
      muDbeta = mu / beta;
      f = f_full();
      DtsAtd = d;
      scal(n, (TR)(beta0 / beta), DtsAtd.data());
      get_gradient(DtsAtd);
      gam = o_gam; Q = 1; C = f;
    }
    // real part only with isreal (complex data, 2d:855-860), then U / scl
    for (i64 i = 0; i < n; ++i) U[i] = finish(U[i], scl);
    std::memcpy(U_out, U.data(), sizeof(TC) * n);
    info->secs = 0;
    info->outer_iters = count_outer;
    info->total_iters = count_total;
    info->rel_chg = RelChgOut;
    if (Ut) {
      V dlt(Ut, Ut + n);
      axpy(n, (TR)-1, U.data(), dlt.data());
      const TS a = snrm2(n, dlt.data()), bb = snrm2(n, Ut);
      info->rel_error = a / bb;
    } else {
      info->rel_error = 0;
    }
  }
};

}  // namespace

/* ---------------------------------------------------------------------- */
/* C interface for the test harness                                        */
/* ---------------------------------------------------------------------- */
extern "C" {

typedef struct tv3o_op tv3o_op;

void tv3o_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

/* Chunk length of the FP64 reductions (default 4096; >= 1).  Changing it
 * changes only the summation order of dot products and norms. */
void tv3o_set_sum_chunk(long long c) { g_chunk = c >= 1 ? c : 1; }

/* Planted mistake k (see g_plant; 0 restores the faithful oracle). */
void tv3o_plant(int k) { g_plant = k; }

/* A x and A^T x products the solver issued since the last reset (the blocked
 * dense A^T A of get_g counts one of each). */
void tv3o_reset_counts(void) { g_nfwd = g_nbwd = 0; }
void tv3o_counts(long long *fwd, long long *bwd) {
  if (fwd) *fwd = g_nfwd;
  if (bwd) *bwd = g_nbwd;
}

int tv3o_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

tv3o_op *tv3o_op_dense(int m, int n, const float *A) {
  Op *op = new (std::nothrow) Op();
  if (!op) return nullptr;
  op->kind = 0; op->m = m; op->n = n; op->A = A; op->nnz = (i64)m * n;
  return (tv3o_op *)op;
}

tv3o_op *tv3o_op_sparse(int m, int n, int nnz, const int *csr_ptr,
                        const int *csr_ind, const float *csr_val,
                        const int *csc_ptr, const int *csc_ind,
                        const float *csc_val) {
  Op *op = new (std::nothrow) Op();
  if (!op) return nullptr;
  op->kind = 1; op->m = m; op->n = n; op->nnz = nnz;
  op->csr_ptr = csr_ptr; op->csr_ind = csr_ind; op->csr_val = csr_val;
  op->csc_ptr = csc_ptr; op->csc_ind = csc_ind; op->csc_val = csc_val;
  return (tv3o_op *)op;
}

/* Complex operators (the complex overloads): entries re, im interleaved as
 * doubles (complex<float> entries widen exactly).  Borrowed. */
tv3o_op *tv3o_op_dense_c(int m, int n, const double *Az) {
  Op *op = new (std::nothrow) Op();
  if (!op) return nullptr;
  op->kind = 0; op->m = m; op->n = n; op->Az = Az; op->nnz = (i64)m * n;
  return (tv3o_op *)op;
}

tv3o_op *tv3o_op_sparse_c(int m, int n, int nnz, const int *csr_ptr,
                          const int *csr_ind, const double *csr_valz,
                          const int *csc_ptr, const int *csc_ind,
                          const double *csc_valz) {
  Op *op = new (std::nothrow) Op();
  if (!op) return nullptr;
  op->kind = 1; op->m = m; op->n = n; op->nnz = nnz;
  op->csr_ptr = csr_ptr; op->csr_ind = csr_ind; op->csr_valz = csr_valz;
  op->csc_ptr = csc_ptr; op->csc_ind = csc_ind; op->csc_valz = csc_valz;
  return (tv3o_op *)op;
}

tv3o_op *tv3o_op_geometry(const tv3_geometry *g) {
  Op *op = new (std::nothrow) Op();
  if (!op) return nullptr;
  build_geometry(*op, *g);
  return (tv3o_op *)op;
}

void tv3o_op_free(tv3o_op *op) { delete (Op *)op; }

long long tv3o_op_nnz(const tv3o_op *op) { return ((const Op *)op)->nnz; }

/* Walk of every output row (rows in order, walks of one row in loop order). */
int tv3o_op_walk(const tv3o_op *h, long long *ptr, int *lin) {
  const Op &op = *(const Op *)h;
  if (op.kind != 2) return 1;
  i64 k = 0;
  ptr[0] = 0;
  for (i64 r = 0; r < op.m; ++r) {
    for (i64 t = op.row_ptr[r]; t < op.row_ptr[r + 1]; ++t) {
      const int w = op.row_walk[t];
      for (i64 j = op.walk_ptr[w]; j < op.walk_ptr[w + 1]; ++j)
        lin[k++] = op.walk_lin[j];
    }
    ptr[r + 1] = k;
  }
  return 0;
}

int tv3o_gemv(const tv3o_op *h, int trans, const float *x, float *y) {
  gemv(*(const Op *)h, trans != 0, x, y);
  return 0;
}

void tv3o_D(int q, int p, int r, const float *U, float *Ux, float *Uy,
            float *Uz) {
  Dop(Dims{q, p, r}, U, Ux, Uy, Uz);
}

void tv3o_Dt(int q, int p, int r, int exact, const float *X, const float *Y,
             const float *Z, float *res) {
  Dtop(Dims{q, p, r}, exact != 0, X, Y, Z, res);
}

double tv3o_shrink(int q, int p, int r, double beta, int iso, int scal_double,
                   const float *Ux, const float *Uy, const float *Uz,
                   const float *sx, const float *sy, const float *sz,
                   float *Wx, float *Wy, float *Wz) {
  Dims dm{q, p, r};
  if (scal_double)
    return (double)(double)shrink<float, double>(dm, beta, iso != 0, Ux, Uy,
                                                 Uz, sx, sy, sz, Wx, Wy, Wz);
  return (double)(float)shrink<float, float>(dm, (float)beta, iso != 0, Ux,
                                             Uy, Uz, sx, sy, sz, Wx, Wy, Wz);
}

double tv3o_dot(long long n, const float *x, const float *y) {
  return ddot(n, x, y);
}

/* U's pixel count is A.n times the number of layers; several layers only
 * for 2-D dense and sparse operators (2d/tval3_cpu_2d.cpp:737-748 forbids
 * them for geometry) */
bool layers_ok(const Op &op, const Dims &dm) {
  if (op.n <= 0 || dm.n() % op.n != 0) return false;
  const i64 L = dm.n() / op.n;
  return L == 1 || (L > 1 && dm.r == 1 && op.kind != 2);
}

int tv3o_solve(const tv3o_op *h, const tv3_options *o, int q, int p, int r,
               const float *b, float *U, const float *Ut, tv3_info *info,
               tv3_trace_rec *trace, int trace_cap) {
  const Op &op = *(const Op *)h;
  const Dims dm{q, p, r};
  if (!layers_ok(op, dm)) return 1;
  if (o->scal_double) {
    Solver<float, double> s(op, *o, dm);
    s.run(b, U, Ut, info, trace, trace_cap);
  } else {
    Solver<float, float> s(op, *o, dm);
    s.run(b, U, Ut, info, trace, trace_cap);
  }
  return 0;
}

/* The reference's double/double instantiation (tval3_cpu_2d(mat<double>&,
 * ...), 2d/tval3_cpu_2d.cpp:26-29): FP64 vectors and scalars, the FP64
 * solve against which reduced-precision GPU results are measured. */
int tv3o_solve_f64(const tv3o_op *h, const tv3_options *o, int q, int p,
                   int r, const double *b, double *U, const double *Ut,
                   tv3_info *info, tv3_trace_rec *trace, int trace_cap) {
  const Op &op = *(const Op *)h;
  const Dims dm{q, p, r};
  if (!layers_ok(op, dm)) return 1;
  Solver<double, double> s(op, *o, dm);
  s.run(b, U, Ut, info, trace, trace_cap);
  return 0;
}

/* get_g_mult's blocked dense product (test entry point): Au = A U, g = the
 * block-accumulated A^T Au. */
int tv3o_AtA_blocked(const tv3o_op *h, const float *U, int submat_mb, float *Au, float *g) {
  const Op &op = *(const Op *)h;
  if (op.kind != 0) return 1;
  dense_AtA_blocked<float>(op, U, Au, g, submat_mb);
  return 0;
}

int tv3o_gemv_f64(const tv3o_op *h, int trans, const double *x, double *y) {
  gemv(*(const Op *)h, trans != 0, x, y);
  return 0;
}

}  // extern "C"

namespace {
/* The complex instantiations (tval3_cpu_2d/3d with mat<complex<float>> /
 * mat<complex<double>>, 2d/tval3_cpu_2d.cpp:31-41, 58-68, 85-95): vectors of
 * interleaved re, im; T_scal = the component type.  Dense and sparse
 * operators need complex entries (tv3o_op_dense_c / _sparse_c); geometry has
 * real weights.  Returns 2 for isotropic TV (no complex form), 3 for an
 * operator without complex entries. */
template <class R>
int solve_c(const tv3o_op *h, const tv3_options *o, int q, int p, int r,
            const R *b, R *U, const R *Ut, tv3_info *info,
            tv3_trace_rec *trace, int trace_cap) {
  typedef std::complex<R> T;
  const Op &op = *(const Op *)h;
  const Dims dm{q, p, r};
  if (!layers_ok(op, dm)) return 1;
  if (o->tv_iso) return 2;
  if (op.kind != 2 && ((op.kind == 0 && !op.Az) || (op.kind == 1 && !op.csr_valz))) return 3;
  Solver<T, R> s(op, *o, dm);
  s.run((const T *)b, (T *)U, (const T *)Ut, info, trace, trace_cap);
  return 0;
}

}  // namespace

extern "C" {

int tv3o_solve_c64(const tv3o_op *h, const tv3_options *o, int q, int p,
                   int r, const float *b, float *U, const float *Ut,
                   tv3_info *info, tv3_trace_rec *trace, int trace_cap) {
  return solve_c<float>(h, o, q, p, r, b, U, Ut, info, trace, trace_cap);
}

int tv3o_solve_c128(const tv3o_op *h, const tv3_options *o, int q, int p,
                    int r, const double *b, double *U, const double *Ut,
                    tv3_info *info, tv3_trace_rec *trace, int trace_cap) {
  return solve_c<double>(h, o, q, p, r, b, U, Ut, info, trace, trace_cap);
}

int tv3o_gemv_c64(const tv3o_op *h, int trans, const float *x, float *y) {
  gemv(*(const Op *)h, trans != 0, (const std::complex<float> *)x, (std::complex<float> *)y);
  return 0;
}

int tv3o_gemv_c128(const tv3o_op *h, int trans, const double *x, double *y) {
  gemv(*(const Op *)h, trans != 0, (const std::complex<double> *)x, (std::complex<double> *)y);
  return 0;
}

/* complex shrinkage (test entry point): returns lam1 (FP64 sum rounded to
 * float) */
double tv3o_shrink_c64(int q, int p, int r, double beta, int wz_fix,
                       const float *Ux, const float *Uy, const float *Uz,
                       const float *sx, const float *sy, const float *sz,
                       float *Wx, float *Wy, float *Wz) {
  typedef std::complex<float> T;
  return (double)(float)shrink_c<float, float>(
      Dims{q, p, r}, (float)beta, wz_fix != 0, (const T *)Ux, (const T *)Uy, (const T *)Uz,
      (const T *)sx, (const T *)sy, (const T *)sz, (T *)Wx, (T *)Wy, (T *)Wz);
}

}  // extern "C"
