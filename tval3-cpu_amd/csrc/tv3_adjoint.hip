// tv3_adjoint.hip — K unrolled TVAL3 inner iterations and their adjoint.
//
// The differentiable map.  Take the first outer iteration of the reference
// solver (2d/tval3_cpu_2d.cpp:752-830, no multiplier update: sigma = delta =
// 0, DtsAtd = 0), run exactly K inner iterations (no stopping test), and hold
// the line-search decisions of a forward solve fixed (the step factor alpha_k
// of min_u, 2d:526-576, and whether step k used the first-step formula: k = 1
// or a 5-strike fallback, which takes alpha = 1).  With W_k = S(D u_k) the
// shrinkage (2d:587-611) the state reduces to u_k and
//   d_k = F(u_k) = Dt(D u_k - S(D u_k)) + m A^T (A u_k - bs),   m = mu / beta
// (get_g + get_gradient, 2d:396-423, 578-585; v - S(v) = clamp(v, -1/beta,
// 1/beta)).  The steps are
//   u_0 = A^T bs,  bs = scl(b) b                              (scaleb, 2d:228-266)
//   u_k = u_{k-1} - alpha_k tau_k d_{k-1}
//   tau_k = <d,d> / (|D d|^2 + m |A d|^2),  d = d_{k-1}        (first step, 2d:431-440)
//   tau_k = |<s,s> / <s,y>|,  s = u_{k-1} - u_{k-2},  y = d_{k-1} - d_{k-2}
//                                                               (BB, 2d:441-451)
// and the output is U = u_K / scl.
//
// The adjoint runs the recursion backwards.  The Jacobian of F is
// J = Dt P D + m A^T A with P the 0/1 mask |D u| <= 1/beta, symmetric, so
// u-bar += J d-bar costs one stencil pass and one A / A^T pair.  The BB and
// first-step quotients are differentiated in closed form.  Forward states are
// checkpointed every `ckpt` iterations (u_k, d_k, u_{k-1}, d_{k-1}) and each
// segment is recomputed before it is reversed.  Vectors are FP32, every
// reduction is an FP64 block sum in fixed order, scalars are FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "tv3_device.h"
#include "tv3_internal.h"

using namespace tv3;

namespace {

size_t ad_align(size_t x) { return (x + 255) & ~(size_t)255; }

// grid-stride loop over the n voxels of a q x p image
#define AD_LOOP(n)                                                              \
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < (n); v += gridDim.x * kBlock)

// fixed-order block sum of NS values: wavefront butterfly, then the four
// wavefronts in order; thread 0 writes part[s * kNblkGroup + blockIdx.x]
// (the tape's launches take at most kNblkGroup blocks)
template <int NS>
__device__ __forceinline__ void ad_block_sum(double (&acc)[NS], double *part) {
  __shared__ double sh[NS][kBlock / 64];
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double t = wave_sum(acc[s]);
    if (ln == 0) sh[s][wv] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double t = 0;
      for (int w = 0; w < kBlock / 64; ++w) t += sh[s][w];
      part[s * kNblkGroup + blockIdx.x] = t;
    }
  }
}

// x- and y-differences at voxel j (periodic, x contiguous)
__device__ __forceinline__ float dxf(const float *u, int j, int q, int p) {
  const Vox w(j, q, p, 1);
  return u[w.xn()] - u[j];
}
__device__ __forceinline__ float dyf(const float *u, int j, int q, int p) {
  const Vox w(j, q, p, 1);
  return u[w.yn()] - u[j];
}
__device__ __forceinline__ float clampb(float v, float t) { return fminf(fmaxf(v, -t), t); }

// out = a x + b y
__global__ __launch_bounds__(kBlock) void k_ad_axpby(int n, float a, const float *x,
                                                     float b, const float *y, float *out) {
  AD_LOOP(n) out[v] = fmaf(a, x[v], y ? b * y[v] : 0.0f);
}

// up to four n-vector copies in one launch (checkpoint saves and segment
// restores); a null source skips its pair
struct Copy4 {
  float *dst[4];
  const float *src[4];
};
__global__ __launch_bounds__(kBlock) void k_ad_copy4(int n, Copy4 c) {
  AD_LOOP(n) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (c.src[i]) c.dst[i][v] = c.src[i][v];
  }
}

// one backward step's carries: ub1 += ubk, db1 += c ubk, then ubk = 0 (the
// two k_ad_axpby launches and the fill they replace, same operations)
__global__ __launch_bounds__(kBlock) void k_ad_carry(int n, float c, float *ubk, float *ub1,
                                                     float *db1) {
  AD_LOOP(n) {
    const float x = ubk[v];
    ub1[v] = fmaf(1.0f, x, 1.0f * ub1[v]);
    db1[v] = fmaf(c, x, 1.0f * db1[v]);
    ubk[v] = 0.0f;
  }
}

// d = Dt(clamp(D u, 1/beta)) + m g
__global__ __launch_bounds__(kBlock) void k_ad_F(int q, int p, const float *u, const float *g,
                                                 float m, float ib, float *d) {
  const int n = q * p;
  AD_LOOP(n) {
    const Vox w(v, q, p, 1);
    const int xp = w.xp(), yp = w.yp();
    float t = clampb(dxf(u, xp, q, p), ib) - clampb(dxf(u, v, q, p), ib);
    t = t + clampb(dyf(u, yp, q, p), ib) - clampb(dyf(u, v, q, p), ib);
    d[v] = fmaf(m, g[v], t);
  }
}

// out = cs out + cx x + cj (Dt (P .* D z) + m h); P from u (|D u| <= 1/beta),
// or P = 1 when u is null.  cs = 0 ignores the old contents of out.
__global__ __launch_bounds__(kBlock) void k_ad_jac(int q, int p, const float *u, float ib,
                                                   const float *z, const float *h, float m,
                                                   float cs, float cx, const float *x,
                                                   float cj, float *out) {
  const int n = q * p;
  AD_LOOP(n) {
    const Vox w(v, q, p, 1);
    const int xp = w.xp(), yp = w.yp();
    auto px = [&](int j) -> float { return (!u || fabsf(dxf(u, j, q, p)) <= ib) ? 1.0f : 0.0f; };
    auto py = [&](int j) -> float { return (!u || fabsf(dyf(u, j, q, p)) <= ib) ? 1.0f : 0.0f; };
    float t = px(xp) * dxf(z, xp, q, p) - px(v) * dxf(z, v, q, p);
    t = t + py(yp) * dyf(z, yp, q, p) - py(v) * dyf(z, v, q, p);
    if (h) t = fmaf(m, h[v], t);
    float o = cj * t;
    if (x) o = fmaf(cx, x[v], o);
    if (cs != 0.0f) o = fmaf(cs, out[v], o);
    out[v] = o;
  }
}

// sum x y  (FP64)
__global__ __launch_bounds__(kBlock) void k_ad_dot(int n, const float *x, const float *y,
                                                   double *part) {
  double acc[1] = {0};
  AD_LOOP(n) acc[0] += (double)x[v] * (double)y[v];
  ad_block_sum<1>(acc, part);
}

// sum |D d|^2
__global__ __launch_bounds__(kBlock) void k_ad_dnorm(int q, int p, const float *d, double *part) {
  double acc[1] = {0};
  AD_LOOP(q * p) {
    const double a = dxf(d, v, q, p), b = dyf(d, v, q, p);
    acc[0] += a * a + b * b;
  }
  ad_block_sum<1>(acc, part);
}

// s = u1 - u0, y = d1 - d0: slot 0 sum s s, slot 1 sum s y
__global__ __launch_bounds__(kBlock) void k_ad_bb(int n, const float *u1, const float *u0,
                                                  const float *d1, const float *d0, double *part) {
  double acc[2] = {0, 0};
  AD_LOOP(n) {
    const double s = (double)u1[v] - (double)u0[v], y = (double)d1[v] - (double)d0[v];
    acc[0] += s * s;
    acc[1] += s * y;
  }
  ad_block_sum<2>(acc, part);
}

// adjoint of tau = |<s,s>/<s,y>|: gs = a1 s + a2 y, gy = a2 s;
// ub1 += gs, ub0 -= gs, db1 += gy, db0 -= gy
__global__ __launch_bounds__(kBlock) void k_ad_bbadj(int n, const float *u1, const float *u0,
                                                     const float *d1, const float *d0,
                                                     double a1, double a2, float *ub1, float *ub0,
                                                     float *db1, float *db0) {
  AD_LOOP(n) {
    const double s = (double)u1[v] - (double)u0[v], y = (double)d1[v] - (double)d0[v];
    const float gs = (float)(a1 * s + a2 * y), gy = (float)(a2 * s);
    ub1[v] += gs;
    ub0[v] -= gs;
    db1[v] += gy;
    db0[v] -= gy;
  }
}

}  // namespace

struct tv3_tape {
  tv3_ctx *ctx = nullptr;
  const tv3_plan *A = nullptr;
  int q = 0, p = 0, n = 0, m = 0, K = 0, ck = 1;
  double mu = 0, beta = 0, mr = 0, scl = 1;
  int scl_case = 0, imax = 0, imin = 0;  // scl = th / |b[imax] - b[imin]| when scl_case != 0
  double th = 1, bdif_signed = 0;
  std::vector<double> alpha, tau, tN, tDn, tAd, tc;  // per step (index k - 1)
  std::vector<int> first;
  // device memory
  void *mem = nullptr;
  GroupDesc *G = nullptr;
  GroupDesc gh{};  // host copy of *G (the instance is always live here)
  InstScal *S = nullptr;
  double *part = nullptr, *dense = nullptr;
  double *hpart = nullptr;  // pinned host copy of part (allocated at the first sum)
  hipEvent_t hev = nullptr;  // marks the copy into hpart (polled, see wait_copy)
  float *bs = nullptr, *ckp = nullptr;      // ckp: checkpoints, 4 n-vectors each
  float *su = nullptr, *sd = nullptr, *sr = nullptr;  // segment states (ck + 2)
  float *tn1 = nullptr, *tn2 = nullptr, *tm1 = nullptr, *tm2 = nullptr;
  float *ub = nullptr, *dbar = nullptr;     // adjoint ring: 3 n-vectors each
  float *bsbar = nullptr;
  int nblk = 1;
  hipStream_t st = nullptr;

  float *ck_vec(int c, int which) { return ckp + ((size_t)c * 4 + which) * n; }
  float *seg_u(int j) { return su + (size_t)j * n; }
  float *seg_d(int j) { return sd + (size_t)j * n; }
  float *seg_r(int j) { return sr + (size_t)j * m; }
  float *ubar(int k) { return ub + (size_t)(((k % 3) + 3) % 3) * n; }
  float *dbr(int k) { return dbar + (size_t)(((k % 3) + 3) % 3) * n; }

  dim3 grid() const { return dim3(nblk); }

  // device-to-device copies and fills on the tape's stream; the first failure
  // is kept for err()
  hipError_t async_err = hipSuccess;
  void keep(hipError_t e) {
    if (e != hipSuccess && async_err == hipSuccess) async_err = e;
  }
  void cp(float *dst, const float *src, size_t count) {
    keep(hipMemcpyAsync(dst, src, sizeof(float) * count, hipMemcpyDeviceToDevice, st));
  }
  void zero(float *dst, size_t count) { keep(hipMemsetAsync(dst, 0, sizeof(float) * count, st)); }

  int err() {
    const hipError_t e = async_err;
    async_err = hipSuccess;
    if (e != hipSuccess) return hip_check(ctx, e, "adjoint copy");
    return hip_check(ctx, hipGetLastError(), "adjoint launch");
  }

  void fwdA(const float *x, float *y) {
    prof_begin(ctx, kProfGemvN);
    launch_fwd<1>(st, G, 1, m, A->dev.kind, A->dev.dtype, x, y, S, 0);
    prof_end(ctx);
    prof_count_host(ctx, gh, 0, 4);
  }
  void bwdA(const float *x, float *y) {
    prof_begin(ctx, kProfGemvT);
    launch_bwd<1>(st, G, 1, n, A->dev.kind, A->dev.dtype, x, y, S, 0, nullptr, (float *)dense, 0);
    prof_end(ctx);
    prof_count_host(ctx, gh, 1, 4);
  }
  // waits for the tape's stream to pass the copy into hpart: an event after
  // it is polled (a stream synchronization parks the host thread and costs
  // its wake-up on every dot product); TV3_SPIN=0, or no completion within
  // two seconds, synchronizes the stream instead
  int wait_copy() {
    static const bool spin = [] {
      const char *e = std::getenv("TV3_SPIN");
      return !(e && e[0] == '0');
    }();
    if (!spin) return hip_check(ctx, hipStreamSynchronize(st), "partials sync");
    int rc = TV3_OK;
    if (!hev) rc = hip_check(ctx, hipEventCreateWithFlags(&hev, hipEventDisableTiming), "partials event");
    if (!rc) rc = hip_check(ctx, hipEventRecord(hev, st), "partials event");
    if (rc) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spins = 1;; ++spins) {
      const hipError_t e = hipEventQuery(hev);
      if (e == hipSuccess) return TV3_OK;
      if (e != hipErrorNotReady) return hip_check(ctx, e, "partials event");
      if ((spins & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2))
        return hip_check(ctx, hipStreamSynchronize(st), "partials sync");
      __builtin_ia32_pause();
    }
  }

  // the block partials of `slots` sums, added on the host in block order;
  // the copy lands in pinned memory (a pageable target is staged through a
  // bounce buffer on every call)
  int host_sum(int slots, double *out) {
    if (!hpart) {
      void *h = nullptr;
      int rc0 = hip_check(ctx, hipHostMalloc(&h, sizeof(double) * 2 * kNblkGroup, hipHostMallocDefault),
                          "partials hipHostMalloc");
      if (rc0) return rc0;
      hpart = (double *)h;
    }
    const size_t cnt = (size_t)(slots - 1) * kNblkGroup + nblk;
    int rc = hip_check(ctx, hipMemcpyAsync(hpart, part, sizeof(double) * cnt, hipMemcpyDeviceToHost, st),
                       "partials copy");
    if (!rc) rc = wait_copy();
    if (rc) return rc;
    for (int s = 0; s < slots; ++s) {
      double t = 0;
      for (int i = 0; i < nblk; ++i) t += hpart[(size_t)s * kNblkGroup + i];
      out[s] = t;
    }
    return TV3_OK;
  }
  int dot(const float *x, const float *y, double *out) {
    hipLaunchKernelGGL(k_ad_dot, grid(), dim3(kBlock), 0, st, n, x, y, part);
    int rc = err();
    return rc ? rc : host_sum(1, out);
  }
  int dotm(const float *x, const float *y, double *out) {
    const int nb = nblk;
    nblk = (int)std::min<long long>(kNblkGroup, (m + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_ad_dot, grid(), dim3(kBlock), 0, st, m, x, y, part);
    int rc = err();
    if (!rc) rc = host_sum(1, out);
    nblk = nb;
    return rc;
  }

  // d = F(u); r = A u - bs kept in r
  int F(const float *u, float *d, float *r) {
    fwdA(u, r);
    hipLaunchKernelGGL(k_ad_axpby, dim3(std::min(kNblkGroup, (m + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                       m, 1.0f, r, -1.0f, bs, r);
    bwdA(r, tn1);
    hipLaunchKernelGGL(k_ad_F, grid(), dim3(kBlock), 0, st, q, p, u, tn1, (float)mr,
                       (float)(1.0 / beta), d);
    return err();
  }

  // tau of step k (1-based) from u_{k-1}, u_{k-2}, d_{k-1}, d_{k-2}; also
  // records the quotient's parts for the adjoint
  int step_tau(int k, const float *u1, const float *u0, const float *d1, const float *d0) {
    const int i = k - 1;
    if (first[i]) {
      double N, Dd, Ad;
      int rc = dot(d1, d1, &N);
      if (rc) return rc;
      hipLaunchKernelGGL(k_ad_dnorm, grid(), dim3(kBlock), 0, st, q, p, d1, part);
      if ((rc = err()) || (rc = host_sum(1, &Dd))) return rc;
      fwdA(d1, tm1);
      if ((rc = err()) || (rc = dotm(tm1, tm1, &Ad))) return rc;
      const double Dn = Dd + mr * Ad;
      tN[i] = N; tDn[i] = Dn; tAd[i] = Ad;
      tau[i] = std::fabs(N / Dn);
    } else {
      hipLaunchKernelGGL(k_ad_bb, grid(), dim3(kBlock), 0, st, n, u1, u0, d1, d0, part);
      double sv[2];
      int rc = err();
      if (rc || (rc = host_sum(2, sv))) return rc;
      tN[i] = sv[0]; tc[i] = sv[1];
      tau[i] = std::fabs(sv[0] / sv[1]);
    }
    return TV3_OK;
  }
};

namespace {

int make_tape(tv3_ctx *ctx, const tv3_plan *A, const tv3_options *o, int q, int p, int K,
              int ck, tv3_tape **out) {
  tv3_tape *t = new (std::nothrow) tv3_tape();
  if (!t) return set_err(ctx, TV3_ERR_ALLOC, "tape allocation");
  t->ctx = ctx; t->A = A; t->q = q; t->p = p; t->n = q * p; t->m = A->dev.m;
  t->K = K; t->ck = ck; t->mu = o->mu; t->beta = o->beta; t->mr = o->mu / o->beta;
  t->st = ctx->stream;
  t->nblk = (int)std::min<long long>(kNblkGroup, (t->n + kBlock - 1) / kBlock);
  const int ncp = K / ck + 1;
  const size_t n = t->n, m = t->m;
  const size_t dense = A->dev.kind == kDense ? 16 * n : 1;
  const size_t bytes = ad_align(sizeof(GroupDesc)) + ad_align(sizeof(InstScal)) +
                       ad_align(sizeof(double) * 2 * kNblkGroup) + ad_align(sizeof(double) * dense) +
                       ad_align(sizeof(float) * m) * 4 + ad_align(sizeof(float) * n * 4 * ncp) +
                       ad_align(sizeof(float) * n * (ck + 2)) * 2 + ad_align(sizeof(float) * m * (ck + 2)) +
                       ad_align(sizeof(float) * n) * 8;
  int rc = hip_check(ctx, hipMalloc(&t->mem, bytes), "tape hipMalloc");
  if (rc) { delete t; return rc; }
  char *c = (char *)t->mem;
  auto take = [&](size_t b) { char *r = c; c += ad_align(b); return (void *)r; };
  t->G = (GroupDesc *)take(sizeof(GroupDesc));
  t->S = (InstScal *)take(sizeof(InstScal));
  t->part = (double *)take(sizeof(double) * 2 * kNblkGroup);
  t->dense = (double *)take(sizeof(double) * dense);
  t->bs = (float *)take(sizeof(float) * m);
  t->tm1 = (float *)take(sizeof(float) * m);
  t->tm2 = (float *)take(sizeof(float) * m);
  t->bsbar = (float *)take(sizeof(float) * m);
  t->ckp = (float *)take(sizeof(float) * n * 4 * ncp);
  t->su = (float *)take(sizeof(float) * n * (ck + 2));
  t->sd = (float *)take(sizeof(float) * n * (ck + 2));
  t->sr = (float *)take(sizeof(float) * m * (ck + 2));
  t->tn1 = (float *)take(sizeof(float) * n);
  t->tn2 = (float *)take(sizeof(float) * n);
  t->ub = (float *)take(sizeof(float) * n * 3);
  t->dbar = (float *)take(sizeof(float) * n * 3);
  GroupDesc gd{};
  gd.q = q; gd.p = p; gd.r = 1; gd.n = t->n; gd.m = t->m;
  gd.plan = A->dev;
  t->gh = gd;
  InstScal s{};
  s.flags = F_LIVE | F_ACTIVE;
  rc = hip_check(ctx, hipMemcpyAsync(t->G, &gd, sizeof(gd), hipMemcpyHostToDevice, t->st), "copy");
  if (!rc) rc = hip_check(ctx, hipMemcpyAsync(t->S, &s, sizeof(s), hipMemcpyHostToDevice, t->st), "copy");
  if (!rc) rc = hip_check(ctx, hipStreamSynchronize(t->st), "copy sync");
  if (rc) { (void)hipFree(t->mem); delete t; return rc; }
  *out = t;
  return TV3_OK;
}

}  // namespace

extern "C" {

int tv3_unrolled_forward(tv3_ctx *ctx, const tv3_plan *A, const tv3_options *opts,
                         int dim_x, int dim_y, int K, int ckpt, const double *alpha,
                         const int *first_step, const float *b, float *U, tv3_tape **tape) {
  if (!ctx) return TV3_ERR_ARG;
  if (!A || !opts || !alpha || !first_step || !b || !U || !tape)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: null argument");
  if (!plan_takes_f32(A->dev))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: the unrolled solver takes float-valued plans");
  if (dim_x <= 0 || dim_y <= 0 || (long long)dim_x * dim_y != A->dev.n)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: dim_x * dim_y must equal the plan's columns");
  if (K < 1 || ckpt < 1 || ckpt > K)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: need K >= 1 and 1 <= ckpt <= K");
  if (!first_step[0])
    return set_err(ctx, TV3_ERR_ARG, "Argument error: step 1 takes the first-step formula");
  if (!(opts->mu > 0) || !(opts->beta > 0))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: mu and beta must be positive");
  *tape = nullptr;
  int rc = hip_check(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  tv3_tape *t = nullptr;
  if ((rc = make_tape(ctx, A, opts, dim_x, dim_y, K, ckpt, &t))) return rc;
  t->alpha.assign(alpha, alpha + K);
  t->first.assign(first_step, first_step + K);
  t->tau.assign(K, 0); t->tN.assign(K, 0); t->tDn.assign(K, 0); t->tAd.assign(K, 0); t->tc.assign(K, 0);
  const int n = t->n, m = t->m;
  // scaleb (2d:228-266): first entries of smallest / largest magnitude
  std::vector<float> hb(m);
  rc = hip_check(ctx, hipMemcpyAsync(hb.data(), b, sizeof(float) * m, hipMemcpyDeviceToHost, t->st), "b copy");
  if (!rc) rc = hip_check(ctx, hipStreamSynchronize(t->st), "b sync");
  if (rc) { tv3_tape_destroy(t); return rc; }
  int imin = 0, imax = 0;
  for (int i = 0; i < m; ++i) {
    if (std::fabs(hb[i]) < std::fabs(hb[imin])) imin = i;
    if (std::fabs(hb[i]) > std::fabs(hb[imax])) imax = i;
  }
  const double bd = (double)hb[imax] - (double)hb[imin], bdif = std::fabs(bd);
  t->imax = imax; t->imin = imin; t->bdif_signed = bd;
  if (m > 0 && bdif < 0.5) { t->scl_case = 1; t->th = 0.5; t->scl = 0.5 / bdif; }
  else if (m > 0 && bdif > 1.5) { t->scl_case = 2; t->th = 1.5; t->scl = 1.5 / bdif; }
  hipLaunchKernelGGL(k_ad_axpby, dim3(std::min(kNblkGroup, (m + kBlock - 1) / kBlock)), dim3(kBlock), 0, t->st,
                     m, (float)t->scl, b, 0.0f, nullptr, t->bs);
  // u_0, d_0
  float *u_prev = t->tn2, *d_prev = nullptr;  // u_{k-2}, d_{k-2} (valid from k = 2)
  float *u_cur = t->seg_u(0), *d_cur = t->seg_d(0);
  t->bwdA(t->bs, u_cur);
  if ((rc = t->F(u_cur, d_cur, t->seg_r(0)))) { tv3_tape_destroy(t); return rc; }
  // rolling states in the segment buffers: slots 0, 1, 2 reused
  int slot = 0;
  auto save_ck = [&](int c, const float *u1, const float *d1, const float *u0, const float *d0) {
    const Copy4 cc{{t->ck_vec(c, 0), t->ck_vec(c, 1), t->ck_vec(c, 2), t->ck_vec(c, 3)}, {u1, d1, u0, d0}};
    hipLaunchKernelGGL(k_ad_copy4, t->grid(), dim3(kBlock), 0, t->st, n, cc);
  };
  save_ck(0, u_cur, d_cur, nullptr, nullptr);
  for (int k = 1; k <= K; ++k) {
    if ((rc = t->step_tau(k, u_cur, u_prev, d_cur, d_prev))) { tv3_tape_destroy(t); return rc; }
    const int ns = (slot + 1) % 3;
    float *u_new = t->seg_u(ns), *d_new = t->seg_d(ns);
    const double step = t->alpha[k - 1] * t->tau[k - 1];
    hipLaunchKernelGGL(k_ad_axpby, t->grid(), dim3(kBlock), 0, t->st, n, (float)(-step), d_cur,
                       1.0f, u_cur, u_new);
    if ((rc = t->F(u_new, d_new, t->seg_r(ns)))) { tv3_tape_destroy(t); return rc; }
    u_prev = u_cur; d_prev = d_cur; u_cur = u_new; d_cur = d_new; slot = ns;
    if (k % ckpt == 0 && k < K) save_ck(k / ckpt, u_cur, d_cur, u_prev, d_prev);
  }
  hipLaunchKernelGGL(k_ad_axpby, t->grid(), dim3(kBlock), 0, t->st, n, (float)(1.0 / t->scl), u_cur,
                     0.0f, nullptr, U);
  // keep u_K for the adjoint of the output scaling
  t->cp(t->tn2, u_cur, n);
  if ((rc = t->err()) || (rc = hip_check(ctx, hipStreamSynchronize(t->st), "forward sync"))) {
    tv3_tape_destroy(t);
    return rc;
  }
  *tape = t;
  return TV3_OK;
}

int tv3_unrolled_backward(tv3_ctx *ctx, tv3_tape *t, const float *dU, float *db, double *dmu) {
  if (!ctx || !t || t->ctx != ctx) return TV3_ERR_ARG;
  if (!dU || !db || !dmu) return set_err(ctx, TV3_ERR_ARG, "Argument error: null argument");
  int rc = hip_check(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  const int n = t->n, m = t->m, K = t->K, ck = t->ck;
  const dim3 gm(std::min(kNblkGroup, (m + kBlock - 1) / kBlock));
  const float ib = (float)(1.0 / t->beta);
  const float mr = (float)t->mr;
  double mbar = 0;
  float hv[2];  // host staging of two entries of db; alive until the final sync
  // U = u_K / scl: u_K-bar = dU / scl, scl-bar = -<dU, u_K> / scl^2
  double sclbar;
  if ((rc = t->dot(dU, t->tn2, &sclbar))) return rc;
  sclbar = -sclbar / (t->scl * t->scl);
  t->zero(t->ub, n * 3);
  t->zero(t->dbar, n * 3);
  t->zero(t->bsbar, m);
  hipLaunchKernelGGL(k_ad_axpby, t->grid(), dim3(kBlock), 0, t->st, n, (float)(1.0 / t->scl), dU,
                     0.0f, nullptr, t->ubar(K));
  // adjoint of d_k = F(u_k): u-bar += J d-bar, bs-bar -= m A d-bar, m-bar += <r_k, A d-bar>
  auto F_adj = [&](int k, const float *u, const float *r) -> int {
    float *dbk = t->dbr(k);
    t->fwdA(dbk, t->tm1);  // a = A d-bar
    double ra;
    int e = t->err();
    if (e || (e = t->dotm(r, t->tm1, &ra))) return e;
    mbar += ra;
    hipLaunchKernelGGL(k_ad_axpby, gm, dim3(kBlock), 0, t->st, m, -mr, t->tm1, 1.0f, t->bsbar, t->bsbar);
    t->bwdA(t->tm1, t->tn1);
    hipLaunchKernelGGL(k_ad_jac, t->grid(), dim3(kBlock), 0, t->st, t->q, t->p, u, ib, dbk,
                       t->tn1, mr, 1.0f, 0.0f, nullptr, 1.0f, t->ubar(k));
    t->zero(dbk, n);
    return t->err();
  };
  const int nseg = (K + ck - 1) / ck;
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int k0 = sg * ck, k1 = std::min(K, k0 + ck);
    // recompute u_{k0-1} .. u_{k1}: segment index j = k - k0 + 1
    {
      const bool prev = k0 > 0;
      const Copy4 cc{{t->seg_u(1), t->seg_d(1), t->seg_u(0), t->seg_d(0)},
                     {t->ck_vec(sg, 0), t->ck_vec(sg, 1), prev ? t->ck_vec(sg, 2) : nullptr,
                      prev ? t->ck_vec(sg, 3) : nullptr}};
      hipLaunchKernelGGL(k_ad_copy4, t->grid(), dim3(kBlock), 0, t->st, n, cc);
    }
    // r_{k0} = A u_{k0} - bs
    t->fwdA(t->seg_u(1), t->seg_r(1));
    hipLaunchKernelGGL(k_ad_axpby, gm, dim3(kBlock), 0, t->st, m, 1.0f, t->seg_r(1), -1.0f, t->bs, t->seg_r(1));
    for (int k = k0 + 1; k <= k1; ++k) {
      const int j = k - k0 + 1;
      const double step = t->alpha[k - 1] * t->tau[k - 1];
      hipLaunchKernelGGL(k_ad_axpby, t->grid(), dim3(kBlock), 0, t->st, n, (float)(-step),
                         t->seg_d(j - 1), 1.0f, t->seg_u(j - 1), t->seg_u(j));
      if ((rc = t->F(t->seg_u(j), t->seg_d(j), t->seg_r(j)))) return rc;
    }
    for (int k = k1; k > k0; --k) {
      const int j = k - k0 + 1;
      const float *uk = t->seg_u(j), *u1 = t->seg_u(j - 1), *d1 = t->seg_d(j - 1);
      if ((rc = F_adj(k, uk, t->seg_r(j)))) return rc;
      // u_k = u_{k-1} - alpha tau d_{k-1}
      const int i = k - 1;
      const double al = t->alpha[i], tau = t->tau[i];
      double tbar;
      if ((rc = t->dot(d1, t->ubar(k), &tbar))) return rc;
      tbar *= -al;
      hipLaunchKernelGGL(k_ad_carry, t->grid(), dim3(kBlock), 0, t->st, n, (float)(-al * tau),
                         t->ubar(k), t->ubar(k - 1), t->dbr(k - 1));
      if (t->first[i]) {
        // tau = N / Dn: d-bar += tbar (2/Dn) (d - tau (Dt D d + m A^T A d)); m-bar -= tbar tau |Ad|^2 / Dn
        const double Dn = t->tDn[i];
        t->fwdA(d1, t->tm1);
        t->bwdA(t->tm1, t->tn1);
        const double c1 = 2.0 * tbar / Dn, c2 = c1 * tau;
        hipLaunchKernelGGL(k_ad_jac, t->grid(), dim3(kBlock), 0, t->st, t->q, t->p, nullptr, ib, d1,
                           t->tn1, mr, 1.0f, (float)c1, d1, (float)(-c2), t->dbr(k - 1));
        mbar -= tbar * tau * t->tAd[i] / Dn;
      } else {
        // tau = |a / c|: d tau/ds = (2 s - tau sg y) / |c|, d tau/dy = -tau sg s / |c|
        const double c = t->tc[i], ac = std::fabs(c), sgn = c < 0 ? -1.0 : 1.0;
        const double a1 = 2.0 * tbar / ac, a2 = -tbar * tau * sgn / ac;
        hipLaunchKernelGGL(k_ad_bbadj, t->grid(), dim3(kBlock), 0, t->st, n, u1, t->seg_u(j - 2),
                           d1, t->seg_d(j - 2), a1, a2, t->ubar(k - 1), t->ubar(k - 2),
                           t->dbr(k - 1), t->dbr(k - 2));
      }
      if ((rc = t->err())) return rc;
    }
    if (sg == 0) {
      // d_0 = F(u_0), u_0 = A^T bs
      if ((rc = F_adj(0, t->seg_u(1), t->seg_r(1)))) return rc;
      t->fwdA(t->ubar(0), t->tm1);
      hipLaunchKernelGGL(k_ad_axpby, gm, dim3(kBlock), 0, t->st, m, 1.0f, t->tm1, 1.0f, t->bsbar, t->bsbar);
    }
  }
  // bs = scl b: b-bar = scl bs-bar, scl-bar += <bs-bar, b> (= <bs-bar, bs> / scl)
  double sb;
  if ((rc = t->dotm(t->bsbar, t->bs, &sb))) return rc;
  sclbar += sb / t->scl;
  hipLaunchKernelGGL(k_ad_axpby, gm, dim3(kBlock), 0, t->st, m, (float)t->scl, t->bsbar, 0.0f, nullptr, db);
  if (t->scl_case) {
    // scl = th / |b[imax] - b[imin]|: d scl / d b[imax] = -scl sgn / bdif, d/d b[imin] = +scl sgn / bdif
    const double sgn = t->bdif_signed < 0 ? -1.0 : 1.0, bdif = std::fabs(t->bdif_signed);
    const double gmax = -t->scl * sgn / bdif * sclbar;
    if ((rc = hip_check(ctx, hipMemcpyAsync(&hv[0], db + t->imax, sizeof(float), hipMemcpyDeviceToHost, t->st), "copy")) ||
        (rc = hip_check(ctx, hipMemcpyAsync(&hv[1], db + t->imin, sizeof(float), hipMemcpyDeviceToHost, t->st), "copy")) ||
        (rc = hip_check(ctx, hipStreamSynchronize(t->st), "sync")))
      return rc;
    if (t->imax != t->imin) {  // equal indices give b_dif = 0: scl_case 1 needs b_dif > 0
      hv[0] = (float)((double)hv[0] + gmax);
      hv[1] = (float)((double)hv[1] - gmax);
    }
    if ((rc = hip_check(ctx, hipMemcpyAsync(db + t->imax, &hv[0], sizeof(float), hipMemcpyHostToDevice, t->st), "copy")) ||
        (rc = hip_check(ctx, hipMemcpyAsync(db + t->imin, &hv[1], sizeof(float), hipMemcpyHostToDevice, t->st), "copy")))
      return rc;
  }
  if ((rc = t->err()) || (rc = hip_check(ctx, hipStreamSynchronize(t->st), "backward sync")))
    return rc;
  *dmu = mbar / t->beta;
  return TV3_OK;
}

void tv3_tape_destroy(tv3_tape *t) {
  if (!t) return;
  if (t->mem) {
    (void)hipStreamSynchronize(t->st);
    (void)hipFree(t->mem);
  }
  if (t->hpart) (void)hipHostFree(t->hpart);
  if (t->hev) (void)hipEventDestroy(t->hev);
  delete t;
}

}  // extern "C"
