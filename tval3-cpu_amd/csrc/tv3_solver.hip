// tv3_solver.hip — the TVAL3 inner loop on the GPU.
//
// Restates the reference solver (2d/tval3_cpu_2d.cpp:97-886,
// 3d/tval3_cpu_3d.cpp:155-1378) as batched kernels over groups of W
// lane-interleaved instances.  The reference's ~107 vector passes per inner
// iteration (2-D) become five fused passes:
//   k_tau_first             get_tau's first-iteration dot products
//                           (2d:430-442); the Barzilai-Borwein ones
//                           (2d:443-450) come from k_wstep
//   k_descend               descend (2d:481-493): U = Up - tau d, nonneg, D(U),
//                           dU, and get_g's V, g2 = Dt(V), lam2, lam4
//                           (2d:396-423) in one stencil pass
//   k_mpost                 get_g's Aub, lam3, lam5 and the A^T-skip norms
//   k_bt_*                  update_g (2d:495-524) interpolation + lam2..lam5
//   k_wstep                 update_W (2d:651-671: shrinkage, V, Dt, lam1,
//                           lam2, lam4) fused with s = U - Up, get_gradient
//                           (2d:578-585), the two norms of the stop test
//                           and the next get_tau's <s, s>, <s, y>
//   k_outer1, k_mlp_*       RelChgOut and update_mlp (2d:673-694)
// plus the A / A^T products of tv3_plan.hip.  Every scalar of the reference
// (mu, beta, lam1..5, f, C, Q, gam, tau, alpha, counters) lives per instance
// on the device and is updated by one-block-per-instance fin_* kernels in
// T_scal arithmetic, in the reference's order.  The host reads back only an
// OR of the instance flags at the three data-dependent branch points
// (Armijo test, each backtracking round, inner stop).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "tv3_device.h"
#include "tv3_internal.h"

using namespace tv3;

#ifndef TV3_VEC
#define TV3_VEC float
#endif
// 0: float vectors (this unit); 1: double (tv3_solver_f64.hip); 2: complex
// float (tv3_solver_c64.hip); 3: complex double (tv3_solver_c128.hip)
#ifndef TV3_SOLVER_VARIANT
#define TV3_SOLVER_VARIANT 0
#endif

namespace {

// Vector (T_comp) type of this translation unit: float here; the FP64
// engine (tv3_solver_f64.hip) compiles this file again with double.
typedef TV3_VEC vec_t;
// component type (the type of every T_comp scalar: with complex T_comp those
// stay real-valued)
typedef typename vreal<vec_t>::type vr_t;
constexpr bool kCplx = vreal<vec_t>::cplx;
__device__ __forceinline__ float vfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double vfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ cfloat vfma(float a, cfloat b, cfloat c) {
  return cfloat(fmaf(a, b.x, c.x), fmaf(a, b.y, c.y));
}
__device__ __forceinline__ cdouble vfma(double a, cdouble b, cdouble c) {
  return cdouble(fma(a, b.x, c.x), fma(a, b.y, c.y));
}
__device__ __forceinline__ float vabs(float a) { return fabsf(a); }
__device__ __forceinline__ double vabs(double a) { return fabs(a); }
// |z|: a value on an axis gives its one component's magnitude, else the FP64
// root of the FP64 sum of squares, rounded once
__device__ __forceinline__ double cmod(double a, double b) {
  if (b == 0) return fabs(a);
  if (a == 0) return fabs(b);
  return sqrt(a * a + b * b);
}
__device__ __forceinline__ float vmod(float a) { return fabsf(a); }
__device__ __forceinline__ double vmod(double a) { return fabs(a); }
__device__ __forceinline__ float vmod(cfloat a) { return (float)cmod((double)a.x, (double)a.y); }
__device__ __forceinline__ double vmod(cdouble a) { return cmod(a.x, a.y); }
// terms of the FP64 dot products: real, the exact product; complex
// (conj(a) b), Re = a.x b.x + a.y b.y and Im = a.x b.y - a.y b.x, one FP64
// rounding each
__device__ __forceinline__ double dre(float a, float b) { return (double)a * (double)b; }
__device__ __forceinline__ double dre(double a, double b) { return a * b; }
__device__ __forceinline__ double dre(cfloat a, cfloat b) {
  return (double)a.x * (double)b.x + (double)a.y * (double)b.y;
}
__device__ __forceinline__ double dre(cdouble a, cdouble b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ double dimag(float, float) { return 0.0; }
__device__ __forceinline__ double dimag(double, double) { return 0.0; }
__device__ __forceinline__ double dimag(cfloat a, cfloat b) {
  return (double)a.x * (double)b.y - (double)a.y * (double)b.x;
}
__device__ __forceinline__ double dimag(cdouble a, cdouble b) { return a.x * b.y - a.y * b.x; }
// vector values widened to FP64 (the quadratic line-search coefficients)
__device__ __forceinline__ double to_vd(float a) { return (double)a; }
__device__ __forceinline__ double to_vd(double a) { return a; }
__device__ __forceinline__ cdouble to_vd(cfloat a) { return cdouble((double)a.x, (double)a.y); }
__device__ __forceinline__ cdouble to_vd(cdouble a) { return a; }
// nonneg (2d:455-479): complex keeps the clamped real part only
__device__ __forceinline__ float clamp0(float u) { return (u < 0.0f) ? 0.0f : u; }
__device__ __forceinline__ double clamp0(double u) { return (u < 0.0f) ? 0.0f : u; }
__device__ __forceinline__ cfloat clamp0(cfloat u) { return cfloat((u.x < 0.0f) ? 0.0f : u.x, 0.0f); }
__device__ __forceinline__ cdouble clamp0(cdouble u) { return cdouble((u.x < 0.0) ? 0.0 : u.x, 0.0); }
// fma(alpha, y, x) in precision TS, per component, rounded to the component
template <class TS>
__device__ __forceinline__ float vfma_s(TS alpha, float y, float x) { return (float)fma(alpha, (TS)y, (TS)x); }
template <class TS>
__device__ __forceinline__ double vfma_s(TS alpha, double y, double x) { return (double)fma(alpha, (TS)y, (TS)x); }
template <class TS>
__device__ __forceinline__ cfloat vfma_s(TS alpha, cfloat y, cfloat x) {
  return cfloat((float)fma(alpha, (TS)y.x, (TS)x.x), (float)fma(alpha, (TS)y.y, (TS)x.y));
}
template <class TS>
__device__ __forceinline__ cdouble vfma_s(TS alpha, cdouble y, cdouble x) {
  return cdouble((double)fma(alpha, (TS)y.x, (TS)x.x), (double)fma(alpha, (TS)y.y, (TS)x.y));
}
// U / scl, complex with isreal: the real part only (driver, 2d:855-860)
template <class TS>
__device__ __forceinline__ float vdiv_s(float u, TS scl, int) { return (float)((TS)u / scl); }
template <class TS>
__device__ __forceinline__ double vdiv_s(double u, TS scl, int) { return (double)((TS)u / scl); }
template <class TS>
__device__ __forceinline__ cfloat vdiv_s(cfloat u, TS scl, int isreal) {
  return cfloat((float)((TS)u.x / scl), (float)((TS)(isreal ? 0.0f : u.y) / scl));
}
template <class TS>
__device__ __forceinline__ cdouble vdiv_s(cdouble u, TS scl, int isreal) {
  return cdouble((double)((TS)u.x / scl), (double)((TS)(isreal ? 0.0 : u.y) / scl));
}
// per-instance scalars with T_comp fields of vec_t; the operator launches
// read only their flags and take them through InstScal (same layout)
// the double and complex solvers run one instance per group (W = 1)
constexpr bool kOnlyW1 = TV3_SOLVER_VARIANT != 0;
// the plans the vector type takes, and the entry point named in errors
inline bool plan_ok(const PlanDev &p) {
  return TV3_SOLVER_VARIANT == 1   ? plan_takes_f64(p)
         : TV3_SOLVER_VARIANT == 2 ? plan_takes_c64(p)
         : TV3_SOLVER_VARIANT == 3 ? plan_takes_c128(p)
                                   : plan_takes_f32(p);
}
constexpr const char *kPlanErr =
    TV3_SOLVER_VARIANT == 1   ? "Argument error: tv3_solve_f64 needs an f64 or bf16 dense, double sparse or geometry plan"
    : TV3_SOLVER_VARIANT == 2 ? "Argument error: tv3_solve_c64 needs a c64 dense, complex-float sparse or geometry plan"
    : TV3_SOLVER_VARIANT == 3 ? "Argument error: tv3_solve_c128 needs a c128 dense, complex-double sparse or geometry plan"
                              : "Argument error: an f64 / double-valued or complex plan needs its own entry point "
                                "(tv3_solve_f64, tv3_solve_c64, tv3_solve_c128)";
[[maybe_unused]] constexpr const char *kGemvErr =
    TV3_SOLVER_VARIANT == 1   ? "Argument error: tv3_gemv_f64 needs an f64 or bf16 dense, double sparse or geometry plan"
    : TV3_SOLVER_VARIANT == 2 ? "Argument error: tv3_gemv_c64 needs a c64 dense, complex-float sparse or geometry plan"
                              : "Argument error: tv3_gemv_c128 needs a c128 dense, complex-double sparse or geometry plan";
typedef InstScalT<vr_t> Scal;
inline const InstScal *flags_of(const Scal *s) { return reinterpret_cast<const InstScal *>(s); }

struct Arrs {
  vec_t *U, *Up, *Ux, *Uxp, *Uy, *Uyp, *Uz, *Uzp, *g, *gp, *g2, *g2p;
  vec_t *dU, *dUx, *dUy, *dUz, *dg, *dg2, *d, *DtsAtd, *Wx, *Wy, *Wz;
  vec_t *sx, *sy, *sz, *Atb, *tmpn, *rcdU, *Uout;
  vec_t *b, *Au, *Aup, *dAu, *delta, *lastAu, *tmpm;
};

// threads per block of the one-block-per-instance fin_* kernels
constexpr int kFin = 512;

struct KA {
  const GroupDesc *G;
  Scal *S;
  double *part;
  int nblk;
  uint32_t req, forbid;
};

// With W = 64 every lane of a wavefront handles the same voxel: make the
// index wavefront-uniform so the neighbour arithmetic runs on the scalar unit.
template <int W>
__device__ __forceinline__ int uni(int v) {
  if (W == 64) return __builtin_amdgcn_readfirstlane(v);
  return v;
}

#define NLOOP_BEGIN(W)                                                       \
  const GroupDesc &gd = K.G[blockIdx.y];                                    \
  const int lb = threadIdx.x % (W);                                          \
  const int inst = blockIdx.y * (W) + lb;                                    \
  const Scal &S = K.S[inst];                                             \
  const bool on = sel(S.flags, K.req, K.forbid);                             \
  const long long E = (long long)gd.n * (W);                                 \
  if (on)                                                                    \
    for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < E;  \
         e += (long long)K.nblk * kBlock) {                                  \
      const int v = uni<W>((int)(e / (W)));                                  \
      const long long o = gd.off_n + e;                                      \
      (void)v;

#define MLOOP_BEGIN(W)                                                       \
  const GroupDesc &gd = K.G[blockIdx.y];                                    \
  const int lb = threadIdx.x % (W);                                          \
  const int inst = blockIdx.y * (W) + lb;                                    \
  const Scal &S = K.S[inst];                                             \
  const bool on = sel(S.flags, K.req, K.forbid);                             \
  const long long E = (long long)gd.m * (W);                                 \
  if (on)                                                                    \
    for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < E;  \
         e += (long long)K.nblk * kBlock) {                                  \
      const long long o = gd.off_m + e;

#define LOOP_END }

// element j (a voxel index) of lane lb of this group's n-vector X
#define AT(X, j) (X)[gd.off_n + (long long)(j) * (W) + lb]

// ---------------------------------------------------------------------------
// get_tau (2d:425-453)
// ---------------------------------------------------------------------------
template <int W, bool D3>
__global__ __launch_bounds__(kBlock) void k_tau_first(KA K, Arrs a) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  {
    NLOOP_BEGIN(W)
    const Vox w(v, gd.q, gd.p, gd.r);
    const bool z3 = D3 && gd.r > 1;
    const vec_t d0 = a.d[o];
    const vec_t dxn = AT(a.d, w.xn()), dyn = AT(a.d, w.yn());
    const vec_t dzn = z3 ? AT(a.d, w.zn()) : 0.0f;
    const vec_t dx = dxn - d0;
    const vec_t dy = dyn - d0;
    acc[0] += dre(dx, dx);
    acc[1] += dre(dy, dy);
    if (z3) {
      const vec_t dz = dzn - d0;
      acc[2] += dre(dz, dz);
    }
    acc[3] += dre(d0, d0);
    LOOP_END
  }
  const int sl[4] = {0, 1, 2, 3};
  block_partials<W, 4>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_msq(KA K, const vec_t *x) {
  double acc[1] = {0.0};
  {
    MLOOP_BEGIN(W)
    const vec_t t = x[o];
    acc[0] += dre(t, t);
    LOOP_END
  }
  const int sl[1] = {kSlotM};
  block_partials<W, 1>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

// ---------------------------------------------------------------------------
// descend (2d:481-493) + get_g's n-vector part (2d:396-417)
// ---------------------------------------------------------------------------
// With POLY it also gathers, for the backtracking line search, the
// quadratic coefficients of lam2(alpha) and lam4(alpha): V_dir(alpha) =
// (U_dir_p - W_dir) + alpha dU_dir with dU_dir = U_dir - U_dir_p (slots 7-11).
// D3: the batch holds 3-D volumes (each group still tests its own dim_z).
// k_descend4 (64-lane 2-D float groups, four instances per lane) repeats the
// 2-D arithmetic of this kernel in descend2d_elem; the GPU tests compare the
// two bit for bit.
//
// The vector kernels read every value of an element before they write any:
// the arrays are distinct, but the compiler cannot prove it and would keep
// each load behind the preceding stores, one memory latency at a time.
template <int W, class TS, bool D3, bool POLY>
__global__ __launch_bounds__(kBlock) void k_descend(KA K, Arrs a, int nonneg,
                                                    int exact) {
  double acc[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  {
    NLOOP_BEGIN(W)
    const vr_t tf = (vr_t)(-1 * (TS)S.tau);
    auto un = [&](vec_t dd, vec_t pp) -> vec_t {
      vec_t u = vfma(tf, dd, pp);
      if (nonneg) u = clamp0(u);
      return u;
    };
    auto unj = [&](int j) -> vec_t { return un(AT(a.d, j), AT(a.Up, j)); };
    const Vox w(v, gd.q, gd.p, gd.r);
    const bool z3 = D3 && gd.r > 1;
    const int xn = w.xn(), yn = w.yn(), xp = w.xp(), yp = w.yp();
    // ---- loads
    const vec_t d0 = a.d[o], p0 = a.Up[o];
    const vec_t dxn = AT(a.d, xn), pxn = AT(a.Up, xn);
    const vec_t dyn = AT(a.d, yn), pyn = AT(a.Up, yn);
    const vec_t dxp = AT(a.d, xp), pxp = AT(a.Up, xp);
    const vec_t dyp = AT(a.d, yp), pyp = AT(a.Up, yp);
    const vec_t wx = a.Wx[o], wy = a.Wy[o], sx = a.sx[o], sy = a.sy[o];
    const vec_t wxp = AT(a.Wx, xp), wyp = AT(a.Wy, yp);
    vec_t qx = 0.0f, qy = 0.0f, qz = 0.0f;  // Ux, Uy, Uz of the previous iterate
    if (POLY) {
      qx = a.Uxp[o];
      qy = a.Uyp[o];
    }
    vec_t dzn = 0.0f, pzn = 0.0f, wz = 0.0f, sz = 0.0f;
    if (z3) {
      dzn = AT(a.d, w.zn());
      pzn = AT(a.Up, w.zn());
      wz = a.Wz[o];
      sz = a.sz[o];
      if (POLY) qz = a.Uzp[o];
    }
    // ---- U = Up - tau d, D(U), V = D(U) - W
    const vec_t u0 = un(d0, p0);
    const vec_t ux = un(dxn, pxn) - u0;
    const vec_t uy = un(dyn, pyn) - u0;
    const vec_t uz = z3 ? un(dzn, pzn) - u0 : 0.0f;
    const vec_t Vx0 = ux - wx;
    const vec_t Vy0 = uy - wy;
    const vec_t Vz0 = z3 ? uz - wz : 0.0f;
    acc[0] += dre(Vx0, Vx0);
    acc[1] += dre(Vy0, Vy0);
    acc[3] += dre(sx, Vx0);
    acc[4] += dre(sy, Vy0);
    if (z3) {
      acc[2] += dre(Vz0, Vz0);
      acc[5] += dre(sz, Vz0);
    }
    acc[6] += dre(d0, d0);
    if (POLY) {
      const auto vx = to_vd(qx) - to_vd(wx), dx = to_vd(ux - qx);
      const auto vy = to_vd(qy) - to_vd(wy), dy = to_vd(uy - qy);
      double A2 = dre(vx, vx) + dre(vy, vy), B2 = dre(vx, dx) + dre(vy, dy), C2 = dre(dx, dx) + dre(dy, dy);
      double A4 = dre(to_vd(sx), vx) + dre(to_vd(sy), vy);
      double B4 = dre(to_vd(sx), dx) + dre(to_vd(sy), dy);
      if (z3) {
        const auto vz = to_vd(qz) - to_vd(wz), dz = to_vd(uz - qz);
        A2 += dre(vz, vz); B2 += dre(vz, dz); C2 += dre(dz, dz);
        A4 += dre(to_vd(sz), vz); B4 += dre(to_vd(sz), dz);
      }
      acc[7] += A2; acc[8] += B2; acc[9] += C2; acc[10] += A4; acc[11] += B4;
    }
    // ---- g2 = Dt(V) at v
    const vec_t VXp = (u0 - un(dxp, pxp)) - wxp;
    const vec_t VYp = (u0 - un(dyp, pyp)) - wyp;
    vec_t t = VXp - Vx0;
    if (!z3) {
      t = t + VYp;
      t = t - Vy0;
    } else {
      const int zp = w.zp();
      if (w.v >= w.q && (w.y > 0 || !exact)) {
        vec_t Yb = VYp;
        if (w.y == 0) {  // last row of the previous slice
          const int j = v - w.q;
          Yb = (unj(v - w.s) - unj(j)) - AT(a.Wy, j);
        }
        t = t + Yb;
        t = t - Vy0;
      }
      if (w.y == 0) {
        t = t + VYp;
        t = t - Vy0;
      }
      const vec_t VZp = (u0 - unj(zp)) - AT(a.Wz, zp);
      t = t + VZp;
      t = t - Vz0;
    }
    // ---- stores
    a.U[o] = u0;
    a.Ux[o] = ux;
    a.Uy[o] = uy;
    if (z3) a.Uz[o] = uz;
    a.g2[o] = t;
    LOOP_END
  }
  const int sl[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
  block_partials<W, 12>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

// 16-byte store of a lane's four instances.  all: every instance either
// computed or may be overwritten (finished).  Otherwise the old values of the
// skipping ones are read back and stored unchanged, so that the row is still
// written whole (four 4-byte stores left partly written sectors, which the
// memory completes with a read of its own).
__device__ __forceinline__ void store4_sel(vec_t *p, float4 v, const bool (&on)[4], bool all) {
  if (!all) {
    const float4 old = *reinterpret_cast<const float4 *>(p);
    v.x = on[0] ? v.x : old.x;
    v.y = on[1] ? v.y : old.y;
    v.z = on[2] ? v.z : old.z;
    v.w = on[3] ? v.w : old.w;
  }
  *reinterpret_cast<float4 *>(p) = v;
}

// Block partials of the four-instances-per-lane kernels (lane l holds
// instances 4 (l mod 16) .. +3 of voxel slot l / 16): each wavefront adds
// its four slots in order into its own LDS rows, then the four wavefronts
// are added in order; one partial per computing instance and slot.
template <int NS>
__device__ __forceinline__ void block_partials4(const double (&acc)[4][NS], const int (&slot)[NS],
                                                const KA &K) {
  __shared__ double red[kBlock / 64][NS][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int quad = lane & 15, vs = lane >> 4;
#pragma unroll
  for (int sl = 0; sl < 4; ++sl) {
    if (vs == sl) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          double &r = red[wv][s][4 * quad + k];
          r = sl == 0 ? acc[k][s] : r + acc[k][s];
        }
    }
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NS * 64; t += kBlock) {
    const int s = t / 64, b = t % 64;
    if (!sel(K.S[blockIdx.y * 64 + b].flags, K.req, K.forbid)) continue;
    const double x = ((red[0][s][b] + red[1][s][b]) + red[2][s][b]) + red[3][s][b];
    K.part[((long long)(blockIdx.y * 64 + b) * kNslot + slot[s]) * kNblkMax + blockIdx.x] = x;
  }
}

// One element of k_descend for a 2-D volume (the same operations in the same
// order as its !z3 path): U, Ux, Uy and g2 = Dt(V) at the voxel, and the
// instance's dot-product terms (slots 0, 1, 3, 4, 6 and, with POLY, 7-11).
template <class TS, bool POLY>
__device__ __forceinline__ void descend2d_elem(vr_t tf, int nonneg, vec_t d0, vec_t p0,
                                               vec_t dxn, vec_t pxn, vec_t dyn, vec_t pyn,
                                               vec_t dxp, vec_t pxp, vec_t dyp, vec_t pyp,
                                               vec_t wx, vec_t wy, vec_t sx, vec_t sy,
                                               vec_t wxp, vec_t wyp, vec_t qx, vec_t qy,
                                               double (&acc)[12], vec_t &u0, vec_t &ux,
                                               vec_t &uy, vec_t &t) {
  auto un = [&](vec_t dd, vec_t pp) -> vec_t {
    vec_t u = vfma(tf, dd, pp);
    if (nonneg) u = clamp0(u);
    return u;
  };
  u0 = un(d0, p0);
  ux = un(dxn, pxn) - u0;
  uy = un(dyn, pyn) - u0;
  const vec_t Vx0 = ux - wx;
  const vec_t Vy0 = uy - wy;
  acc[0] += dre(Vx0, Vx0);
  acc[1] += dre(Vy0, Vy0);
  acc[3] += dre(sx, Vx0);
  acc[4] += dre(sy, Vy0);
  acc[6] += dre(d0, d0);
  if (POLY) {
    const auto vx = to_vd(qx) - to_vd(wx), dx = to_vd(ux - qx);
    const auto vy = to_vd(qy) - to_vd(wy), dy = to_vd(uy - qy);
    const double A2 = dre(vx, vx) + dre(vy, vy), B2 = dre(vx, dx) + dre(vy, dy), C2 = dre(dx, dx) + dre(dy, dy);
    const double A4 = dre(to_vd(sx), vx) + dre(to_vd(sy), vy);
    const double B4 = dre(to_vd(sx), dx) + dre(to_vd(sy), dy);
    acc[7] += A2; acc[8] += B2; acc[9] += C2; acc[10] += A4; acc[11] += B4;
  }
  const vec_t VXp = (u0 - un(dxp, pxp)) - wxp;
  const vec_t VYp = (u0 - un(dyp, pyp)) - wyp;
  t = VXp - Vx0;
  t = t + VYp;
  t = t - Vy0;
}

// k_descend for 64-lane groups of 2-D float instances, four instances per
// lane: lane l of a wavefront takes instances 4 (l mod 16) .. +3 of voxel
// slot l / 16, so every array access is one 16-byte load or store (a
// wavefront moves four voxels' 256-byte rows per instruction; with one
// float per lane the stencil's 16-18 loads per voxel were issue-bound).
// The element arithmetic is descend2d_elem's; each instance's per-block
// FP64 partial sums its voxels in a fixed order (rounds, then the four
// slots of a wavefront, then the four wavefronts: block_partials4).
template <class TS, bool POLY>
__global__ __launch_bounds__(kBlock) void k_descend4(KA K, Arrs a, int nonneg) {
  const GroupDesc &gd = K.G[blockIdx.y];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int quad = lane & 15, slot = lane >> 4;
  const int i0 = blockIdx.y * 64 + 4 * quad;
  bool on[4];
  vr_t tf[4];
  bool any = false, all = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const Scal &S = K.S[i0 + k];
    on[k] = sel(S.flags, K.req, K.forbid);
    tf[k] = (vr_t)(-1 * (TS)S.tau);
    any = any || on[k];
    // a skipping instance that has finished (or pads the group) keeps no
    // state here: its result went to Uout when it finished, and no later
    // pass reads its working arrays, so its lane may be overwritten
    all = all && (on[k] || !(S.flags & F_ACTIVE));
  }
  double acc[4][12];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int s = 0; s < 12; ++s) acc[k][s] = 0.0;
  if (any) {
    const int n = gd.n, q = gd.q, p = gd.p;
    // element (v, instance 4 quad + k) of a group vector: row v, float4 quad
    auto ld = [&](const vec_t *X, int v) -> float4 {
      return *reinterpret_cast<const float4 *>(X + gd.off_n + (long long)v * 64 + 4 * quad);
    };
    for (int v = blockIdx.x * 16 + wv * 4 + slot; v < n; v += K.nblk * 16) {
      const Vox w(v, q, p, 1);
      const int xn = w.xn(), yn = w.yn(), xp = w.xp(), yp = w.yp();
      const float4 d0 = ld(a.d, v), p0 = ld(a.Up, v);
      const float4 dxn = ld(a.d, xn), pxn = ld(a.Up, xn);
      const float4 dyn = ld(a.d, yn), pyn = ld(a.Up, yn);
      const float4 dxp = ld(a.d, xp), pxp = ld(a.Up, xp);
      const float4 dyp = ld(a.d, yp), pyp = ld(a.Up, yp);
      const float4 wx = ld(a.Wx, v), wy = ld(a.Wy, v), sx = ld(a.sx, v), sy = ld(a.sy, v);
      const float4 wxp = ld(a.Wx, xp), wyp = ld(a.Wy, yp);
      float4 qx = make_float4(0.0f, 0.0f, 0.0f, 0.0f), qy = qx;
      if (POLY) {
        qx = ld(a.Uxp, v);
        qy = ld(a.Uyp, v);
      }
      float4 u0, ux, uy, t;
#define TV3_D4(k, c)                                                                        \
  descend2d_elem<TS, POLY>(tf[k], nonneg, d0.c, p0.c, dxn.c, pxn.c, dyn.c, pyn.c, dxp.c, pxp.c, \
                           dyp.c, pyp.c, wx.c, wy.c, sx.c, sy.c, wxp.c, wyp.c, qx.c, qy.c,       \
                           acc[k], u0.c, ux.c, uy.c, t.c)
      TV3_D4(0, x);
      TV3_D4(1, y);
      TV3_D4(2, z);
      TV3_D4(3, w);
#undef TV3_D4
      const long long o = gd.off_n + (long long)v * 64 + 4 * quad;
      store4_sel(a.U + o, u0, on, all);
      store4_sel(a.Ux + o, ux, on, all);
      store4_sel(a.Uy + o, uy, on, all);
      store4_sel(a.g2 + o, t, on, all);
    }
  }
  const int sl[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
  block_partials4<12>(acc, sl, K);
}

// get_g's m-vector part: Aub = Au - b, lam3 = |Aub|^2, lam5 = <delta, Aub>,
// and the norms of the A^T-skip test (2d:313-316)
// poly != 0: also the quadratic coefficients of lam3(alpha), lam5(alpha)
// with Au(alpha) = Aup + alpha (Au - Aup) (slots kSlotM+4 .. +8).
template <int W, bool SKIP, bool POLY>
__global__ __launch_bounds__(kBlock) void k_mpost(KA K, Arrs a) {
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  {
    MLOOP_BEGIN(W)
    const vec_t au = a.Au[o];
    const vec_t bb = a.b[o];
    const vec_t dl = a.delta[o];
    const vec_t la = SKIP ? a.lastAu[o] : 0.0f;
    const vec_t aup = POLY ? a.Aup[o] : 0.0f;
    const vec_t aub = au - bb;
    acc[0] += dre(aub, aub);
    acc[1] += dre(dl, aub);
    if (SKIP) {
      const vec_t da = au - la;
      acc[2] += dre(da, da);
      acc[3] += dre(au, au);
    }
    if (POLY) {
      const auto ap = to_vd(aup) - to_vd(bb), da = to_vd(au - aup);
      acc[4] += dre(ap, ap);
      acc[5] += dre(ap, da);
      acc[6] += dre(da, da);
      acc[7] += dre(to_vd(dl), ap);
      acc[8] += dre(to_vd(dl), da);
    }
    LOOP_END
  }
  const int sl[9] = {kSlotM, kSlotM + 1, kSlotM + 2, kSlotM + 3, kSlotM + 4,
                     kSlotM + 5, kSlotM + 6, kSlotM + 7, kSlotM + 8};
  block_partials<W, 9>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_copy_m(KA K, const vec_t *src,
                                                   vec_t *dst) {
  MLOOP_BEGIN(W)
  dst[o] = src[o];
  LOOP_END
}

// g = gp - Atb: get_g with the A^T product skipped keeps the previous g
// (already holding -Atb) and subtracts Atb again (2d:350-353, 2d:400)
template <int W>
__global__ __launch_bounds__(kBlock) void k_skipg(KA K, Arrs a) {
  NLOOP_BEGIN(W)
  a.g[o] = a.gp[o] - a.Atb[o];
  LOOP_END
}

// ---------------------------------------------------------------------------
// min_u backtracking (2d:550-575).  The reference interpolates g2 too
// (2d:509); update_W recomputes g2 from the new W before anything reads it
// (2d:664), so that interpolation is not carried out here.  dU = U - Up is
// formed here, once, from the descent's U (the same float subtraction as
// 2d:542).
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_bt_diff_n(KA K, Arrs a) {
  NLOOP_BEGIN(W)
  const vec_t dg = a.g[o] - a.gp[o];
  const vec_t du = a.U[o] - a.Up[o];
  const vec_t dux = a.Ux[o] - a.Uxp[o];
  const vec_t duy = a.Uy[o] - a.Uyp[o];
  const vec_t duz = gd.r > 1 ? a.Uz[o] - a.Uzp[o] : 0.0f;
  a.dg[o] = dg;
  a.dU[o] = du;
  a.dUx[o] = dux;
  a.dUy[o] = duy;
  if (gd.r > 1) a.dUz[o] = duz;
  LOOP_END
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_bt_diff_m(KA K, Arrs a) {
  MLOOP_BEGIN(W)
  a.dAu[o] = a.Au[o] - a.Aup[o];
  LOOP_END
}

template <class TS>
__device__ __forceinline__ vec_t interp(vec_t x, TS alpha, vec_t y) {
  return vfma_s(alpha, y, x);
}

template <int W, class TS>
__global__ __launch_bounds__(kBlock) void k_bt_upd_n(KA K, Arrs a) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  {
    NLOOP_BEGIN(W)
    const TS al = (TS)S.alpha;
    const bool z3 = gd.r > 1;
    // ---- loads
    const vec_t gp = a.gp[o], dg = a.dg[o];
    const vec_t up = a.Up[o], du = a.dU[o];
    const vec_t uxp = a.Uxp[o], dux = a.dUx[o], uyp = a.Uyp[o], duy = a.dUy[o];
    const vec_t wx = a.Wx[o], wy = a.Wy[o], sx = a.sx[o], sy = a.sy[o];
    vec_t uzp = 0.0f, duz = 0.0f, wz = 0.0f, sz = 0.0f;
    if (z3) {
      uzp = a.Uzp[o];
      duz = a.dUz[o];
      wz = a.Wz[o];
      sz = a.sz[o];
    }
    // ---- the state at alpha, lam2, lam4
    const vec_t ux = interp<TS>(uxp, al, dux);
    const vec_t uy = interp<TS>(uyp, al, duy);
    const vec_t Vx = ux - wx, Vy = uy - wy;
    acc[0] += dre(Vx, Vx);
    acc[1] += dre(Vy, Vy);
    acc[3] += dre(sx, Vx);
    acc[4] += dre(sy, Vy);
    vec_t uz = 0.0f;
    if (z3) {
      uz = interp<TS>(uzp, al, duz);
      const vec_t Vz = uz - wz;
      acc[2] += dre(Vz, Vz);
      acc[5] += dre(sz, Vz);
    }
    // ---- stores
    a.g[o] = interp<TS>(gp, al, dg);
    a.U[o] = interp<TS>(up, al, du);
    a.Ux[o] = ux;
    a.Uy[o] = uy;
    if (z3) a.Uz[o] = uz;
    LOOP_END
  }
  const int sl[6] = {0, 1, 2, 3, 4, 5};
  block_partials<W, 6>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

template <int W, class TS>
__global__ __launch_bounds__(kBlock) void k_bt_upd_m(KA K, Arrs a) {
  double acc[2] = {0, 0};
  {
    MLOOP_BEGIN(W)
    const TS al = (TS)S.alpha;
    const vec_t au = interp<TS>(a.Aup[o], al, a.dAu[o]);
    a.Au[o] = au;
    const vec_t aub = au - a.b[o];
    acc[0] += dre(aub, aub);
    acc[1] += dre(a.delta[o], aub);
    LOOP_END
  }
  const int sl[2] = {kSlotM, kSlotM + 1};
  block_partials<W, 2>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

// The state at the step size chosen by the quadratic line search, built as
// the reference's last backtracking round builds it: dX = X - Xp from the
// full step, X = Xp + alpha dX (update_g, 2d:500-505), then lam2/lam4 and
// lam3/lam5 from the rounded vectors.
template <int W, class TS>
__global__ __launch_bounds__(kBlock) void k_interp_n(KA K, Arrs a) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  {
    NLOOP_BEGIN(W)
    const TS al = (TS)S.alpha;
    const bool z3 = gd.r > 1;
    // ---- loads
    const vec_t gp = a.gp[o], gc = a.g[o];
    const vec_t up = a.Up[o], du = a.U[o] - up;
    const vec_t uxp = a.Uxp[o], uyp = a.Uyp[o], uxc = a.Ux[o], uyc = a.Uy[o];
    const vec_t wx = a.Wx[o], wy = a.Wy[o], sx = a.sx[o], sy = a.sy[o];
    vec_t uzp = 0.0f, uzc = 0.0f, wz = 0.0f, sz = 0.0f;
    if (z3) {
      uzp = a.Uzp[o];
      uzc = a.Uz[o];
      wz = a.Wz[o];
      sz = a.sz[o];
    }
    // ---- the state at alpha, lam2, lam4
    const vec_t ux = interp<TS>(uxp, al, uxc - uxp);
    const vec_t uy = interp<TS>(uyp, al, uyc - uyp);
    const vec_t Vx = ux - wx, Vy = uy - wy;
    acc[0] += dre(Vx, Vx);
    acc[1] += dre(Vy, Vy);
    acc[3] += dre(sx, Vx);
    acc[4] += dre(sy, Vy);
    vec_t uz = 0.0f;
    if (z3) {
      uz = interp<TS>(uzp, al, uzc - uzp);
      const vec_t Vz = uz - wz;
      acc[2] += dre(Vz, Vz);
      acc[5] += dre(sz, Vz);
    }
    // ---- stores
    a.g[o] = interp<TS>(gp, al, gc - gp);
    a.U[o] = interp<TS>(up, al, du);
    a.Ux[o] = ux;
    a.Uy[o] = uy;
    if (z3) a.Uz[o] = uz;
    LOOP_END
  }
  const int sl[6] = {0, 1, 2, 3, 4, 5};
  block_partials<W, 6>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

template <int W, class TS>
__global__ __launch_bounds__(kBlock) void k_interp_m(KA K, Arrs a) {
  double acc[2] = {0, 0};
  {
    MLOOP_BEGIN(W)
    const TS al = (TS)S.alpha;
    const vec_t aup = a.Aup[o];
    const vec_t au = interp<TS>(aup, al, a.Au[o] - aup);
    a.Au[o] = au;
    const vec_t aub = au - a.b[o];
    acc[0] += dre(aub, aub);
    acc[1] += dre(a.delta[o], aub);
    LOOP_END
  }
  const int sl[2] = {kSlotM, kSlotM + 1};
  block_partials<W, 2>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

// ---------------------------------------------------------------------------
// D (2d:268-278) alone, for the initial U = A^T b
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_D(KA K, const vec_t *U, vec_t *Ux,
                                              vec_t *Uy, vec_t *Uz) {
  NLOOP_BEGIN(W)
  const Vox w(v, gd.q, gd.p, gd.r);
  const vec_t u0 = U[o];
  Ux[o] = AT(U, w.xn()) - u0;
  Uy[o] = AT(U, w.yn()) - u0;
  if (gd.r > 1) Uz[o] = AT(U, w.zn()) - u0;
  LOOP_END
}

// ---------------------------------------------------------------------------
// update_W (2d:651-671) + get_gradient + stop-test norms + get_tau's
// <s, s>, <s, y> of the next iteration
// ---------------------------------------------------------------------------
// k_wstep4 (64-lane 2-D float groups after the first iteration, four
// instances per lane) repeats the 2-D arithmetic of this kernel in
// wstep2d_elem; the GPU tests compare the two bit for bit.
template <int W, class TS, bool INIT, bool D3, bool ISO>
__global__ __launch_bounds__(kBlock) void k_wstep(KA K, Arrs a, int exact) {
  // complex: slot kSlotSyIm also takes Im <s, y>
  constexpr int NW = kCplx ? 11 : 10;
  double acc[NW] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  {
    NLOOP_BEGIN(W)
    const TS beta = (TS)S.beta;
    const TS ib = (TS)1 / beta;
    const bool z3 = D3 && gd.r > 1;
    const Vox w(v, gd.q, gd.p, gd.r);
    // D(U) and sigma at one voxel; the anisotropic shrinkage of a neighbour
    // needs one direction only (dir 0, 1, 2; -1: all)
    struct Cell {
      vec_t ux, uy, uz, sx, sy, sz;
    };
    auto cell = [&](int j, int dir) -> Cell {
      const long long oj = gd.off_n + (long long)j * W + lb;
      Cell c{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      if (ISO || dir <= 0) { c.ux = a.Ux[oj]; c.sx = a.sx[oj]; }
      // the reference's complex 3-D Wz takes the y direction's ubar
      if (ISO || dir < 0 || dir == 1 || (kCplx && dir == 2 && !exact)) { c.uy = a.Uy[oj]; c.sy = a.sy[oj]; }
      if (z3 && (ISO || dir < 0 || dir == 2)) { c.uz = a.Uz[oj]; c.sz = a.sz[oj]; }
      return c;
    };
    // shrinkage of one voxel; returns the T_scal lam1 term
    auto shr = [&](const Cell &c, vec_t &wx, vec_t &wy, vec_t &wz) -> TS {
      if constexpr (kCplx) {
        const auto bx = cshrink_bar(c.ux, c.sx, beta, ib), by = cshrink_bar(c.uy, c.sy, beta, ib);
        wx = cshrink_w(bx.w, bx.bar, bx.mag);
        wy = cshrink_w(by.w, by.bar, by.mag);
        TS t = bx.w + by.w;
        wz = 0.0f;
        if (z3) {
          const auto bz = cshrink_bar(c.uz, c.sz, beta, ib);
          wz = cshrink_w(bz.w, exact ? bz.bar : by.bar, bz.mag);
          t = t + bz.w;
        }
        return t;
      } else {
        if (ISO)
          return shrink_iso<TS>(c.ux, c.uy, c.uz, c.sx, c.sy, c.sz, z3, beta, ib, wx, wy, wz);
        TS tx, ty, tz = 0;
        wx = shrink1<TS>(c.ux, c.sx, beta, ib, tx);
        wy = shrink1<TS>(c.uy, c.sy, beta, ib, ty);
        wz = z3 ? shrink1<TS>(c.uz, c.sz, beta, ib, tz) : 0.0f;
        TS t = tx + ty;
        if (z3) t = t + tz;
        return t;
      }
    };
    // ---- loads
    const int xp = w.xp(), yp = w.yp();
    const Cell c0 = cell(v, -1);
    const Cell cx = cell(xp, 0);
    const Cell cy = cell(yp, 1);
    Cell cz{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, cj = cz;
    if (z3) {
      cz = cell(w.zp(), 2);
      // the y-term from the last row of the previous slice
      if (w.v >= w.q && w.y == 0 && !exact) cj = cell(v - w.q, 1);
    }
    vec_t up = 0.0f, uc = 0.0f, gc = 0.0f, gpv = 0.0f, g2pv = 0.0f, dts = 0.0f;
    if (!INIT) {
      up = a.Up[o];
      uc = a.U[o];
      gc = a.g[o];
      gpv = a.gp[o];
      g2pv = a.g2p[o];
      dts = a.DtsAtd[o];
    }
    // ---- W = shrink(D(U) + sigma/beta), V = D(U) - W, lam1, lam2, lam4
    vec_t wx0, wy0, wz0;
    const TS l1 = shr(c0, wx0, wy0, wz0);
    acc[0] += (double)l1;
    const vec_t Vx0 = c0.ux - wx0, Vy0 = c0.uy - wy0;
    const vec_t Vz0 = z3 ? c0.uz - wz0 : 0.0f;
    acc[1] += dre(Vx0, Vx0);
    acc[2] += dre(Vy0, Vy0);
    acc[4] += dre(c0.sx, Vx0);
    acc[5] += dre(c0.sy, Vy0);
    if (z3) {
      acc[3] += dre(Vz0, Vz0);
      acc[6] += dre(c0.sz, Vz0);
    }
    // ---- g2 = Dt(V) at v
    vec_t tx, ty, tz;
    shr(cx, tx, ty, tz);
    const vec_t VXp = cx.ux - tx;
    shr(cy, tx, ty, tz);
    const vec_t VYp = cy.uy - ty;
    vec_t t = VXp - Vx0;
    if (!z3) {
      t = t + VYp;
      t = t - Vy0;
    } else {
      if (w.v >= w.q && (w.y > 0 || !exact)) {
        vec_t Yb = VYp;
        if (w.y == 0) {
          shr(cj, tx, ty, tz);
          Yb = cj.uy - ty;
        }
        t = t + Yb;
        t = t - Vy0;
      }
      if (w.y == 0) {
        t = t + VYp;
        t = t - Vy0;
      }
      shr(cz, tx, ty, tz);
      const vec_t VZp = cz.uz - tz;
      t = t + VZp;
      t = t - Vz0;
    }
    vec_t dn = 0.0f;
    if (!INIT) {
      const vec_t uu = uc - up;
      acc[7] += dre(up, up);
      acc[8] += dre(uu, uu);
      const vr_t mdb = (vr_t)S.muDbeta;
      dn = vfma(mdb, gc, dts + t);
      // the next iteration's Barzilai-Borwein step (get_tau, 2d:443-450):
      // s = U - Up, y = (g2 - g2p) + (mu/beta) (g - gp).  No kernel between
      // here and the next get_tau changes these vectors or mu/beta of an
      // instance that is not restarting its inner loop (which takes the
      // first-iteration formula instead), so the dot products are formed
      // here; slot 8 holds <s, s>.
      const vec_t y = vfma(mdb, gc - gpv, t - g2pv);
      acc[9] += dre(uu, y);
      if constexpr (kCplx) acc[NW - 1] += dimag(uu, y);
    }
    // ---- stores
    a.Wx[o] = wx0;
    a.Wy[o] = wy0;
    if (z3) a.Wz[o] = wz0;
    a.g2[o] = t;
    if (!INIT) a.d[o] = dn;
    LOOP_END
  }
  if constexpr (kCplx) {
    const int sl[11] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, kSlotSyIm};
    block_partials<W, 11>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
  } else {
    const int sl[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
    block_partials<W, 10>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
  }
}

// One element of k_wstep for a 2-D float volume after the first iteration
// (the same operations in the same order as its !INIT, !z3 path): W at the
// voxel, g2 = Dt(V), d, and the instance's terms (slots 0-2, 4, 5, 7-9).
// c0, cx, cy: (Ux, sx, Uy, sy) at the voxel, at x - 1 and at y - 1 (the
// anisotropic shrinkage reads one direction of cx and cy only).
template <class TS, bool ISO>
__device__ __forceinline__ void wstep2d_elem(TS beta, TS ib, float mdb, const float (&c0)[4],
                                             const float (&cx)[4], const float (&cy)[4], float up,
                                             float uc, float gc, float gpv, float g2pv, float dts,
                                             double (&acc)[10], float &wx0, float &wy0, float &t,
                                             float &dn) {
  // shrinkage of one cell (ux, sx, uy, sy); returns the lam1 term
  auto shr = [&](const float (&c)[4], float &wx, float &wy) -> TS {
    float wz;
    if (ISO) return shrink_iso<TS>(c[0], c[2], 0.0f, c[1], c[3], 0.0f, false, beta, ib, wx, wy, wz);
    TS tx, ty;
    wx = shrink1<TS>(c[0], c[1], beta, ib, tx);
    wy = shrink1<TS>(c[2], c[3], beta, ib, ty);
    return tx + ty;
  };
  const TS l1 = shr(c0, wx0, wy0);
  acc[0] += (double)l1;
  const float Vx0 = c0[0] - wx0, Vy0 = c0[2] - wy0;
  acc[1] += dre(Vx0, Vx0);
  acc[2] += dre(Vy0, Vy0);
  acc[4] += dre(c0[1], Vx0);
  acc[5] += dre(c0[3], Vy0);
  float tx, ty;
  shr(cx, tx, ty);
  const float VXp = cx[0] - tx;
  shr(cy, tx, ty);
  const float VYp = cy[2] - ty;
  t = VXp - Vx0;
  t = t + VYp;
  t = t - Vy0;
  const float uu = uc - up;
  acc[7] += dre(up, up);
  acc[8] += dre(uu, uu);
  dn = vfma(mdb, gc, dts + t);
  const float y = vfma(mdb, gc - gpv, t - g2pv);
  acc[9] += dre(uu, y);
}

// k_wstep after the first iteration for 64-lane groups of 2-D float
// instances, four instances per lane (the layout of k_descend4).
template <class TS, bool ISO>
__global__ __launch_bounds__(kBlock) void k_wstep4(KA K, Arrs a) {
  const GroupDesc &gd = K.G[blockIdx.y];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int quad = lane & 15, slot = lane >> 4;
  const int i0 = blockIdx.y * 64 + 4 * quad;
  bool on[4];
  TS beta[4], ib[4];
  float mdb[4];
  bool any = false, all = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const Scal &S = K.S[i0 + k];
    on[k] = sel(S.flags, K.req, K.forbid);
    beta[k] = (TS)S.beta;
    ib[k] = (TS)1 / beta[k];
    mdb[k] = (float)(vr_t)S.muDbeta;
    any = any || on[k];
    // a skipping instance that has finished (or pads the group) keeps no
    // state here: its result went to Uout when it finished, and no later
    // pass reads its working arrays, so its lane may be overwritten
    all = all && (on[k] || !(S.flags & F_ACTIVE));
  }
  double acc[4][10];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int s = 0; s < 10; ++s) acc[k][s] = 0.0;
  if (any) {
    const int n = gd.n, q = gd.q, p = gd.p;
    auto ld = [&](const vec_t *X, int v) -> float4 {
      return *reinterpret_cast<const float4 *>(X + gd.off_n + (long long)v * 64 + 4 * quad);
    };
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int v = blockIdx.x * 16 + wv * 4 + slot; v < n; v += K.nblk * 16) {
      const Vox w(v, q, p, 1);
      const int xp = w.xp(), yp = w.yp();
      const float4 ux0 = ld(a.Ux, v), sx0 = ld(a.sx, v), uy0 = ld(a.Uy, v), sy0 = ld(a.sy, v);
      const float4 uxx = ld(a.Ux, xp), sxx = ld(a.sx, xp);
      const float4 uyx = ISO ? ld(a.Uy, xp) : z4, syx = ISO ? ld(a.sy, xp) : z4;
      const float4 uxy = ISO ? ld(a.Ux, yp) : z4, sxy = ISO ? ld(a.sx, yp) : z4;
      const float4 uyy = ld(a.Uy, yp), syy = ld(a.sy, yp);
      const float4 up = ld(a.Up, v), uc = ld(a.U, v), gc = ld(a.g, v);
      const float4 gpv = ld(a.gp, v), g2pv = ld(a.g2p, v), dts = ld(a.DtsAtd, v);
      float4 wx, wy, t, dn;
#define TV3_W4(k, c)                                                                        \
  {                                                                                         \
    const float c0[4] = {ux0.c, sx0.c, uy0.c, sy0.c};                                       \
    const float cx[4] = {uxx.c, sxx.c, uyx.c, syx.c};                                       \
    const float cy[4] = {uxy.c, sxy.c, uyy.c, syy.c};                                       \
    wstep2d_elem<TS, ISO>(beta[k], ib[k], mdb[k], c0, cx, cy, up.c, uc.c, gc.c, gpv.c,      \
                          g2pv.c, dts.c, acc[k], wx.c, wy.c, t.c, dn.c);                    \
  }
      TV3_W4(0, x);
      TV3_W4(1, y);
      TV3_W4(2, z);
      TV3_W4(3, w);
#undef TV3_W4
      const long long o = gd.off_n + (long long)v * 64 + 4 * quad;
      store4_sel(a.Wx + o, wx, on, all);
      store4_sel(a.Wy + o, wy, on, all);
      store4_sel(a.g2 + o, t, on, all);
      store4_sel(a.d + o, dn, on, all);
    }
  }
  const int sl[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  block_partials4<10>(acc, sl, K);
}

// d = DtsAtd + g2 + muDbeta*g, with DtsAtd = (beta0/beta) d first when
// scale != 0 (driver, 2d:845-849)
template <int W>
__global__ __launch_bounds__(kBlock) void k_grad(KA K, Arrs a, int scale) {
  NLOOP_BEGIN(W)
  vec_t dts = a.DtsAtd[o];
  const vec_t dc = scale ? a.d[o] : 0.0f;
  const vec_t gc = a.g[o], g2c = a.g2[o];
  if (scale) {
    dts = (vr_t)S.ratio * dc;
    a.DtsAtd[o] = dts;
  }
  a.d[o] = vfma((vr_t)S.muDbeta, gc, dts + g2c);
  LOOP_END
}

// ---------------------------------------------------------------------------
// outer iteration (2d:820-852)
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(kBlock) void k_outer1(KA K, Arrs a) {
  double acc[1] = {0};
  {
    NLOOP_BEGIN(W)
    const vec_t u = a.U[o];
    const vec_t df = u - a.rcdU[o];
    acc[0] += dre(df, df);
    a.rcdU[o] = u;
    LOOP_END
  }
  const int sl[1] = {0};
  block_partials<W, 1>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_mlp_n(KA K, Arrs a) {
  double acc[3] = {0, 0, 0};
  {
    NLOOP_BEGIN(W)
    const vr_t nb = (vr_t)(-1 * S.beta);
    const bool z3 = gd.r > 1;
    const vec_t Vx = a.Ux[o] - a.Wx[o], Vy = a.Uy[o] - a.Wy[o];
    const vec_t sx = vfma(nb, Vx, a.sx[o]), sy = vfma(nb, Vy, a.sy[o]);
    vec_t Vz = 0.0f, sz = 0.0f;
    if (z3) {
      Vz = a.Uz[o] - a.Wz[o];
      sz = vfma(nb, Vz, a.sz[o]);
    }
    a.sx[o] = sx;
    a.sy[o] = sy;
    acc[0] += dre(sx, Vx);
    acc[1] += dre(sy, Vy);
    if (z3) {
      a.sz[o] = sz;
      acc[2] += dre(sz, Vz);
    }
    LOOP_END
  }
  const int sl[3] = {0, 1, 2};
  block_partials<W, 3>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_mlp_m(KA K, Arrs a) {
  double acc[1] = {0};
  {
    MLOOP_BEGIN(W)
    const vr_t nm = (vr_t)(-1 * S.mu);
    const vec_t aub = a.Au[o] - a.b[o];
    const vec_t dl = vfma(nm, aub, a.delta[o]);
    a.delta[o] = dl;
    acc[0] += dre(dl, aub);
    LOOP_END
  }
  const int sl[1] = {kSlotM};
  block_partials<W, 1>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

// U_out = U / scl (driver, 2d:856-860)
template <int W, class TS>
__global__ __launch_bounds__(kBlock) void k_finish(KA K, Arrs a, int isreal) {
  NLOOP_BEGIN(W)
  a.Uout[o] = vdiv_s(a.U[o], (TS)S.scl, isreal);
  LOOP_END
}

// |Ut - U|^2 and |Ut|^2 for info.rel_error (2d:869-874)
template <int W>
__global__ __launch_bounds__(kBlock) void k_relerr(KA K, const vec_t *Ut,
                                                   const vec_t *U) {
  double acc[2] = {0, 0};
  {
    NLOOP_BEGIN(W)
    const vec_t t = Ut[o];
    const vec_t df = t - U[o];
    acc[0] += dre(df, df);
    acc[1] += dre(t, t);
    LOOP_END
  }
  const int sl[2] = {0, 1};
  block_partials<W, 2>(acc, sl, K.part, blockIdx.y * W, blockIdx.x, K.S, K.req, K.forbid);
}

// ---------------------------------------------------------------------------
// scaleb (2d:228-266): first index of min |b| and of max |b|
// ---------------------------------------------------------------------------
struct MinMax {
  vr_t mn, mx;
  int imn, imx;
};

__device__ __forceinline__ void mm_merge(MinMax &a, const MinMax &b) {
  if (b.mn < a.mn || (b.mn == a.mn && b.imn < a.imn)) { a.mn = b.mn; a.imn = b.imn; }
  if (b.mx > a.mx || (b.mx == a.mx && b.imx < a.imx)) { a.mx = b.mx; a.imx = b.imx; }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_absminmax(KA K, const vec_t *bv,
                                                      MinMax *mmpart) {
  const GroupDesc &gd = K.G[blockIdx.y];
  const int lb = threadIdx.x % W;
  const int inst = blockIdx.y * W + lb;
  MinMax r{INFINITY, -INFINITY, 0x7fffffff, 0x7fffffff};
  const long long E = (long long)gd.m * W;
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < E;
       e += (long long)K.nblk * kBlock) {
    const vr_t t = vmod(bv[gd.off_m + e]);
    const int i = (int)(e / W);
    MinMax c{t, t, i, i};
    mm_merge(r, c);
  }
  __shared__ MinMax red[kBlock];
  red[threadIdx.x] = r;
  __syncthreads();
  for (int s = kBlock / 2; s >= W; s >>= 1) {
    if ((int)threadIdx.x < s) mm_merge(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if ((int)threadIdx.x < W)
    mmpart[(long long)inst * kNblkMax + blockIdx.x] = red[threadIdx.x];
}

// b = (vec_t)scl * b (cblas_scal takes a vec_t alpha)
template <int W>
__global__ __launch_bounds__(kBlock) void k_scale_m(KA K, vec_t *x) {
  MLOOP_BEGIN(W)
  x[o] = (vr_t)S.scl * x[o];
  LOOP_END
}

// U = rcdU = x (initial guess U = A^T b, 2d:775, 787)
template <int W>
__global__ __launch_bounds__(kBlock) void k_init_U(KA K, Arrs a) {
  NLOOP_BEGIN(W)
  const vec_t t = a.Atb[o];
  a.U[o] = t;
  a.rcdU[o] = t;
  LOOP_END
}

// pack caller vectors (one pointer per lane, NULL = zeros) into the
// lane-interleaved layout, and back
template <int W>
__global__ __launch_bounds__(kBlock) void k_pack(const GroupDesc *G,
                                                 const vec_t *const *src,
                                                 vec_t *dst, int use_m) {
  const GroupDesc &gd = G[blockIdx.y];
  const long long len = use_m ? gd.m : gd.n;
  const long long off = use_m ? gd.off_m : gd.off_n;
  const long long E = len * W;
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < E;
       e += (long long)gridDim.x * kBlock) {
    const int lb = (int)(e % W);
    const vec_t *p = src[blockIdx.y * W + lb];
    dst[off + e] = p ? p[e / W] : 0.0f;
  }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_unpack(const GroupDesc *G,
                                                   const vec_t *src,
                                                   vec_t *const *dst) {
  const GroupDesc &gd = G[blockIdx.y];
  const long long E = (long long)gd.n * W;
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < E;
       e += (long long)gridDim.x * kBlock) {
    const int lb = (int)(e % W);
    vec_t *p = dst[blockIdx.y * W + lb];
    if (p) p[e / W] = src[gd.off_n + e];
  }
}

// ===========================================================================
// per-instance scalar control (one 256-thread block per instance)
// ===========================================================================
struct FA {
  Scal *S;
  const GroupDesc *G;
  const double *part;
  int nblk_n, nblk_m;
  int W;
  uint32_t req, forbid;
  unsigned long long *fout;  // device view of the mapped host flag words (or NULL)
  uint32_t epoch;             // this launch's number, written with the flags
  tv3_trace_rec *trace;
  int trace_cap;
  OptDev O;
};

// Sums of NS partial-sum slots of one instance by a kFin-thread block, in one
// pass: thread t adds partials t, t+kFin, ...; the DPP tree of wave_sum;
// the wavefronts in order.  Every thread receives the results.
// (one-block-per-instance kernels run with kFin threads)
template <int NS>
__device__ __forceinline__ void fin_sums(const FA &F, int inst,
                                         const int (&slots)[NS],
                                         double (&out)[NS]) {
  __shared__ double red[kFin / 64][NS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double x[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) x[k] = 0.0;
  // one partial per thread and slot: all loads are issued before any
  // reduction.  The loads are unconditional (every slot holds kNblkMax
  // entries; those past the slot's block count are stale and not added):
  // a load under a per-slot branch would wait for its value inside the
  // branch, one memory round trip per slot.
  const int nbmax = F.nblk_n > F.nblk_m ? F.nblk_n : F.nblk_m;
  for (int i = threadIdx.x; i < nbmax; i += kFin) {
    double t[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) t[k] = F.part[((long long)inst * kNslot + slots[k]) * kNblkMax + i];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const int nb = slots[k] >= kSlotM ? F.nblk_m : F.nblk_n;
      x[k] = i < nb ? x[k] + t[k] : x[k];
    }
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const double t = wave_sum(x[k]);
    if (lane == 0) red[wv][k] = t;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    double t = red[0][k];
    for (int w = 1; w < kFin / 64; ++w) t += red[w][k];
    out[k] = t;
  }
  __syncthreads();
}

__device__ __forceinline__ void put_flags(const FA &F, int inst, uint32_t f) {
  if (F.fout) {
    // one 64-bit store: a host that reads the launch's number reads its
    // flags.  A system-scope store goes to host memory by itself; the host
    // reads nothing else the kernel wrote, so no fence (a write-back of the
    // whole L2 that the block would wait for) orders other stores before it.
    __hip_atomic_store(F.fout + inst, ((unsigned long long)F.epoch << 32) | f,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// f = lam1 + beta/2*lam2 + mu/2*lam3 - lam4 - lam5 (2d:422, 523, 842)
template <class TS>
__device__ __forceinline__ vr_t f_full(const Scal &s) {
  const TS lam1 = (TS)s.lam1, lam2 = (TS)s.lam2, lam3 = (TS)s.lam3;
  const TS beta = (TS)s.beta, mu = (TS)s.mu;
  return (vr_t)((((lam1 + beta / 2 * lam2) + mu / 2 * lam3) - (TS)s.lam4) -
                 (TS)s.lam5);
}

// the Armijo test of min_u (2d:548, 558): |f| > |C - alpha*c_armij|
template <class TS>
__device__ __forceinline__ bool armijo_fails(const Scal &s) {
  return (TS)vabs(s.f) > fabs((TS)s.C - (TS)s.alpha * (TS)s.c_armij);
}

#define FIN_BEGIN                                                      \
  const int inst = blockIdx.x;                                         \
  Scal &s = F.S[inst];                                             \
  const uint32_t fl0 = s.flags;                                        \
  const bool on = sel(fl0, F.req, F.forbid);                           \
  if (!on) {                                                           \
    if (threadIdx.x == 0) put_flags(F, inst, fl0);                     \
    return;                                                            \
  }                                                                    \
  const GroupDesc &gd = F.G[inst / F.W];                               \
  (void)gd;

#define FIN_END                                                        \
  if (threadIdx.x == 0) put_flags(F, inst, s.flags);

// lam2 and lam4 from three direction dot products (get_g / update_g /
// update_W: 2d:408-411)
template <class TS>
__device__ __forceinline__ void set_lam24(Scal &s, bool z3, double x2,
                                          double y2, double z2, double x4,
                                          double y4, double z4) {
  TS lam2 = (TS)(vr_t)x2;
  lam2 += (TS)(vr_t)y2;
  if (z3) lam2 += (TS)(vr_t)z2;
  vr_t lam4 = (vr_t)x4;
  lam4 += (vr_t)y4;
  if (z3) lam4 += (vr_t)z4;
  s.lam2 = (double)lam2;
  s.lam4 = lam4;
}

constexpr uint32_t kStepFlags =
    F_MLP | F_JUST_DONE | F_INNER_DONE | F_BT | F_FB | F_BWD | F_SKIPG;

// tau of get_tau (2d:425-453).  Normal mode (req without F_FB): also clears
// the per-step flags of every instance.  Fallback mode (req = F_FB): the
// first-iteration formula for the instances in the fallback.
template <class TS>
__global__ __launch_bounds__(kFin) void fin_tau(FA F) {
  const int inst = blockIdx.x;
  Scal &s = F.S[inst];
  const uint32_t fl = s.flags;
  const bool fb = (F.req & F_FB) != 0;
  if (fb ? !(fl & F_FB) : !sel(fl, F.req, F.forbid)) {
    if (threadIdx.x == 0 && !fb) {
      s.flags = fl & ~kStepFlags;
      put_flags(F, inst, s.flags);
    }
    return;
  }
  const GroupDesc &gd = F.G[inst / F.W];
  const bool first = fb || (fl & F_FST);
  double r[5] = {0, 0, 0, 0, 0};
  if (first) {
    const int sl[5] = {0, 1, 2, 3, kSlotM};
    fin_sums<5>(F, inst, sl, r);
  } else {
    // <s, s> and <s, y> (complex: also Im <s, y>), formed by the previous
    // update_W pass
    if constexpr (kCplx) {
      const int sl[3] = {8, 9, kSlotSyIm};
      double r2[3];
      fin_sums<3>(F, inst, sl, r2);
      r[0] = r2[0];
      r[1] = r2[1];
      r[2] = r2[2];
    } else {
      const int sl[2] = {8, 9};
      double r2[2];
      fin_sums<2>(F, inst, sl, r2);
      r[0] = r2[0];
      r[1] = r2[1];
    }
  }
  if (threadIdx.x == 0) {
    TS tau;
    if (first) {
      vr_t dDd = (vr_t)r[0] + (vr_t)r[1];
      if (gd.r > 1) dDd = dDd + (vr_t)r[2];
      const vr_t dd = (vr_t)r[3], Add = (vr_t)r[4];
      tau = fabs((TS)dd / ((TS)dDd + (TS)s.muDbeta * (TS)Add));
    } else {
      const vr_t ss = (vr_t)r[0], sy = (vr_t)r[1];
      if constexpr (kCplx) {
        // abs(ss / sy) of complex values (ss real-valued): |ss| / |sy| in
        // FP64, rounded to the component type
        const double den = cmod((double)sy, (double)(vr_t)r[2]);
        tau = (TS)(vr_t)(fabs((double)ss) / den);
      } else {
        tau = (TS)vabs(ss / sy);
      }
    }
    s.tau = (double)tau;
    if (!fb) {
      s.flags = fl & ~kStepFlags;
      s.cnt = 0;
      s.fallback = 0;
      s.alpha = 1.0;
    }
    put_flags(F, inst, s.flags);
  }
}

// A^T-skipping state machine of get_g_mult (2d:300-357).  force: the
// initial get_g and the 5-strike fallback.
template <class TS>
__global__ __launch_bounds__(kFin) void fin_skip(FA F, int force) {
  FIN_BEGIN
  const int sl[2] = {kSlotM + 2, kSlotM + 3};
  double r[2];
  fin_sums<2>(F, inst, sl, r);
  if (threadIdx.x == 0) {
    const TS nrm_d = (TS)(vr_t)sqrt(r[0]), nrm_a = (TS)(vr_t)sqrt(r[1]);
    const TS relChange = nrm_d / nrm_a;
    uint32_t fl = s.flags & ~(F_BWD | F_SKIPG);
    TS csv = (TS)s.currSkipValue;
    if (force) {
      fl |= F_BWD; csv = 0;
    } else if (s.currInitIteration < F.O.numInitIterations) {
      fl |= F_BWD; s.currInitIteration++; csv = 0;
    } else if (csv < 1) {
      fl |= F_BWD; csv += (TS)F.O.skipMulRatio;
    } else if (relChange > (TS)F.O.maxRelativeChange) {
      fl |= F_BWD;
    } else {
      fl |= F_SKIPG; csv -= 1;
    }
    s.currSkipValue = (double)csv;
    s.flags = fl;
  }
  FIN_END
}

// End of descend/get_g: lam2..5 and f.  mode 0: the Armijo test of min_u
// (2d:543-549) and, if it fails, the backtracking step size: with
// ls_exact = 0 the five candidate steps alpha = gamma^k are evaluated from
// the quadratic f(alpha) whose coefficients the descent pass gathered (the
// state for the chosen alpha is then rebuilt exactly by k_interp_*), with
// ls_exact = 1 one round at a time by k_bt_upd_* and fin_lams.  mode 1:
// after the 5-strike fallback.  mode 2: the initial get_g.
template <class TS>
__global__ __launch_bounds__(kFin) void fin_descend(FA F, int mode) {
  FIN_BEGIN
  const bool z3 = gd.r > 1;
  const bool poly = mode == 0 && !F.O.ls_exact;
  double r[19];
  if (poly) {
    const int sl[19] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, kSlotM, kSlotM + 1,
                        kSlotM + 4, kSlotM + 5, kSlotM + 6, kSlotM + 7, kSlotM + 8};
    fin_sums<19>(F, inst, sl, r);
  } else {
    const int sl[9] = {0, 1, 2, 3, 4, 5, 6, kSlotM, kSlotM + 1};
    double q[9];
    fin_sums<9>(F, inst, sl, q);
    for (int k = 0; k < 7; ++k) r[k] = q[k];
    r[12] = q[7];
    r[13] = q[8];
  }
  if (threadIdx.x == 0) {
    set_lam24<TS>(s, z3, r[0], r[1], r[2], r[3], r[4], r[5]);
    s.lam3 = (double)(TS)(vr_t)r[12];
    s.lam5 = (vr_t)r[13];
    s.f = f_full<TS>(s);
    if (mode == 0) {
      s.c_armij = (double)((TS)(vr_t)r[6] * (TS)s.tau * (TS)F.O.c * (TS)s.beta);
      s.alpha = 1.0;
      if (armijo_fails<TS>(s)) {
        s.cnt = 0;
        if (!poly) {
          s.flags |= F_BT;
          s.alpha = (double)((TS)s.alpha * (TS)F.O.gamma);
        } else {
          // lam2(a) = A2 + 2a B2 + a^2 C2, lam4(a) = A4 + a B4, same for
          // lam3 / lam5 (U_dir(a) = U_dir_p + a dU_dir, Au(a) likewise)
          const double A2 = r[7], B2 = r[8], C2 = r[9], A4 = r[10], B4 = r[11];
          const double A3 = r[14], B3 = r[15], C3 = r[16], A5 = r[17], B5 = r[18];
          const double lam1 = s.lam1, hb = (double)((TS)s.beta / 2),
                       hm = (double)((TS)s.mu / 2);
          TS al = (TS)s.alpha;
          bool ok = false;
          for (int k = 1; k <= 5 && !ok; ++k) {
            al = al * (TS)F.O.gamma;
            const double a = (double)al;
            const double f = lam1 + hb * (A2 + a * (2 * B2 + a * C2)) +
                             hm * (A3 + a * (2 * B3 + a * C3)) -
                             (A4 + a * B4) - (A5 + a * B5);
            s.cnt = k;
            ok = !(fabs(f) > fabs((double)s.C - a * s.c_armij));
          }
          s.alpha = (double)al;
          if (ok) {
            s.flags |= F_BT;  // rebuild the state at alpha (k_interp_*)
          } else {
            s.flags |= F_FB;
            s.gam = (double)((TS)s.gam * (TS)F.O.rate_gam);
          }
        }
      }
    } else if (mode == 1) {
      s.flags &= ~F_FB;
      s.fallback = 1;
    }
  }
  FIN_END
}

// lam2..lam5 and f after an interpolation of the state (update_g,
// 2d:495-524).  mode 0: one reference backtracking round (2d:555-574);
// mode 1: the final interpolation of the quadratic line search.
template <class TS>
__global__ __launch_bounds__(kFin) void fin_lams(FA F, int mode) {
  FIN_BEGIN
  const bool z3 = gd.r > 1;
  const int sl[8] = {0, 1, 2, 3, 4, 5, kSlotM, kSlotM + 1};
  double r[8];
  fin_sums<8>(F, inst, sl, r);
  if (threadIdx.x == 0) {
    set_lam24<TS>(s, z3, r[0], r[1], r[2], r[3], r[4], r[5]);
    s.lam3 = (double)(TS)(vr_t)r[6];
    s.lam5 = (vr_t)r[7];
    s.f = f_full<TS>(s);
    if (mode == 1) {
      s.flags &= ~F_BT;
    } else {
      s.cnt++;
      if (!armijo_fails<TS>(s)) {
        s.flags &= ~F_BT;
      } else if (s.cnt == 5) {
        s.flags = (s.flags & ~F_BT) | F_FB;
        s.gam = (double)((TS)s.gam * (TS)F.O.rate_gam);
      } else {
        s.alpha = (double)((TS)s.alpha * (TS)F.O.gamma);
      }
    }
  }
  FIN_END
}

// update_W's scalars, the nonmonotone reference update, the inner stop test
// (2d:653-670, 799-817).  init: the initial shrinkage + get_g (lam1, lam2,
// lam4 only).
template <class TS>
__global__ __launch_bounds__(kFin) void fin_wstep(FA F, int init) {
  FIN_BEGIN
  const bool z3 = gd.r > 1;
  const int sl[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
  double r[9];
  fin_sums<9>(F, inst, sl, r);
  if (threadIdx.x == 0) {
    const TS beta = (TS)s.beta;
    if (init) {
      s.lam1 = (double)(TS)r[0];
      set_lam24<TS>(s, z3, r[1], r[2], r[3], r[4], r[5], r[6]);
    } else {
      s.f = (vr_t)((TS)s.f - (((TS)s.lam1 + beta / 2 * (TS)s.lam2) - (TS)s.lam4));
      s.lam1 = (double)(TS)r[0];
      set_lam24<TS>(s, z3, r[1], r[2], r[3], r[4], r[5], r[6]);
      s.f = (vr_t)((TS)s.f + (((TS)s.lam1 + beta / 2 * (TS)s.lam2) - (TS)s.lam4));
      const TS gam = (TS)s.gam;
      const TS Qp = (TS)s.Q;
      const TS Q = gam * Qp + 1;
      s.Q = (double)Q;
      s.C = (vr_t)((gam * Qp * (TS)s.C + (TS)s.f) / Q);
      s.count_total++;
      const TS nrmup = (TS)(vr_t)sqrt(r[7]);
      const TS relchg = (TS)(vr_t)sqrt(r[8]) / nrmup;
      s.nrmup = (double)nrmup;
      s.RelChg = (double)relchg;
      if (F.trace && s.count_total <= F.trace_cap) {
        tv3_trace_rec &t = F.trace[(long long)inst * F.trace_cap + s.count_total - 1];
        t.iter = s.count_total; t.outer = s.count_outer;
        t.backtracks = s.cnt; t.fallback = s.fallback;
        t.tau = s.tau; t.alpha = s.alpha; t.f = s.f; t.C = s.C;
        t.rel_chg = s.RelChg; t.lam1 = s.lam1; t.lam2 = s.lam2;
        t.lam3 = s.lam3; t.lam4 = s.lam4; t.lam5 = s.lam5;
        t.mu = s.mu; t.beta = s.beta;
      }
      const bool fst = (s.flags & F_FST) != 0;
      if ((relchg < (TS)F.O.tol_inn && !fst) || s.count_total >= F.O.maxit)
        s.flags |= F_INNER_DONE;
      else
        s.flags &= ~F_FST;
    }
  }
  FIN_END
}

// initial f (end of the first get_g) and the nonmonotone reference values
template <class TS>
__global__ __launch_bounds__(kFin) void fin_init2(FA F) {
  FIN_BEGIN
  const int sl[2] = {kSlotM, kSlotM + 1};
  double r[2];
  fin_sums<2>(F, inst, sl, r);
  if (threadIdx.x == 0) {
    s.lam3 = (double)(TS)(vr_t)r[0];
    s.lam5 = (vr_t)r[1];
    s.f = f_full<TS>(s);
    s.Q = 1.0;
    s.C = s.f;
    if (!(0 < F.O.maxcnt && 0 < F.O.maxit))
      s.flags = (s.flags & ~F_ACTIVE) | F_DONE | F_JUST_DONE;
  }
  FIN_END
}

// scaleb (2d:228-266) and the penalty clamps (2d:776-778)
template <class TS>
__global__ __launch_bounds__(kBlock) void fin_scaleb(FA F, const MinMax *mmp,
                                                     const vec_t *bv) {
  FIN_BEGIN
  __shared__ MinMax red[kBlock];
  MinMax r{INFINITY, -INFINITY, 0x7fffffff, 0x7fffffff};
  for (int i = threadIdx.x; i < F.nblk_m; i += kBlock)
    mm_merge(r, mmp[(long long)inst * kNblkMax + i]);
  red[threadIdx.x] = r;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) mm_merge(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    TS scl = 1;
    const int m = gd.m;
    if (m > 0) {
      const MinMax &mm = red[0];
      const int lbx = inst % F.W;
      // no finite candidate (all NaN / inf): keep b[0], as the reference's
      // scan starts from it
      const int imn = (mm.imn >= 0 && mm.imn < m) ? mm.imn : 0;
      const int imx = (mm.imx >= 0 && mm.imx < m) ? mm.imx : 0;
      const vec_t bmin = bv[gd.off_m + (long long)imn * F.W + lbx];
      const vec_t bmax = bv[gd.off_m + (long long)imx * F.W + lbx];
      const TS b_dif = vmod(bmax - bmin);
      const TS th1 = (TS)0.5, th2 = (TS)1.5;
      if (b_dif < th1)
        scl = th1 / b_dif;
      else if (b_dif > th2)
        scl = th2 / b_dif;
    }
    s.scl = (double)scl;
    TS mu = (TS)F.O.mu0, beta = (TS)F.O.beta0;
    const TS muf = (TS)F.O.mu, betaf = (TS)F.O.beta;
    if (mu > muf) mu = muf;
    if (beta > betaf) beta = betaf;
    s.mu = (double)mu;
    s.beta = (double)beta;
    s.muDbeta = (double)(mu / beta);
    s.gam = (double)(TS)F.O.gam;
  }
  FIN_END
}

// RelChgOut and the outer stop test (2d:820-828)
template <class TS>
__global__ __launch_bounds__(kFin) void fin_outer1(FA F) {
  FIN_BEGIN
  const int sl[1] = {0};
  double r[1];
  fin_sums<1>(F, inst, sl, r);
  if (threadIdx.x == 0) {
    s.count_outer++;
    const TS rr = (TS)(vr_t)sqrt(r[0]) / (TS)s.nrmup;
    s.RelChgOut = (double)rr;
    uint32_t fl = s.flags & ~F_INNER_DONE;
    if (rr < (TS)F.O.tol)
      fl = (fl & ~F_ACTIVE) | F_DONE | F_JUST_DONE;
    else
      fl |= F_MLP;
    s.flags = fl;
  }
  FIN_END
}

// update_mlp's scalars and the continuation step (2d:673-694, 831-851)
template <class TS>
__global__ __launch_bounds__(kFin) void fin_mlp(FA F) {
  FIN_BEGIN
  const bool z3 = gd.r > 1;
  const int sl[4] = {0, 1, 2, kSlotM};
  double r[4];
  fin_sums<4>(F, inst, sl, r);
  if (threadIdx.x == 0) {
    s.f = s.f + (s.lam4 + s.lam5);
    vr_t lam4 = (vr_t)r[0];
    lam4 += (vr_t)r[1];
    if (z3) lam4 += (vr_t)r[2];
    s.lam4 = lam4;
    s.lam5 = (vr_t)r[3];
    s.f = s.f - (s.lam4 + s.lam5);
    const TS beta0 = (TS)s.beta;
    TS beta = beta0 * (TS)F.O.rate_cnt;
    TS mu = (TS)s.mu * (TS)F.O.rate_cnt;
    if (beta > (TS)F.O.beta) beta = (TS)F.O.beta;
    if (mu > (TS)F.O.mu) mu = (TS)F.O.mu;
    s.beta0 = (double)beta0;
    s.beta = (double)beta;
    s.mu = (double)mu;
    s.muDbeta = (double)(mu / beta);
    s.f = f_full<TS>(s);
    s.ratio = (double)(vr_t)(beta0 / beta);
    s.gam = (double)(TS)F.O.gam;
    s.Q = 1.0;
    s.C = s.f;
    if (s.count_outer < F.O.maxcnt && s.count_total < F.O.maxit)
      s.flags |= F_FST;
    else
      s.flags = (s.flags & ~F_ACTIVE) | F_DONE | F_JUST_DONE;
  }
  FIN_END
}

// end of a step: drops the per-step flags of every instance and publishes
// the flags
__global__ __launch_bounds__(kFin) void fin_refresh(FA F) {
  if (threadIdx.x != 0) return;
  Scal &s = F.S[blockIdx.x];
  s.flags &= ~kStepFlags;
  put_flags(F, blockIdx.x, s.flags);
}

template <class TS>
__global__ __launch_bounds__(kFin) void fin_relerr(FA F) {
  FIN_BEGIN
  const int sl[2] = {0, 1};
  double r[2];
  fin_sums<2>(F, inst, sl, r);
  if (threadIdx.x == 0)
    s.rel_error = (double)((TS)(vr_t)sqrt(r[0]) / (TS)(vr_t)sqrt(r[1]));
  FIN_END
}

__global__ void k_init_scal(Scal *S, int B, int Bpad, unsigned long long *fout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Bpad) return;
  Scal z;
  memset(&z, 0, sizeof(z));
  z.flags = i < B ? (F_LIVE | F_ACTIVE | F_FST) : 0u;
  z.alpha = 1.0;
  z.Q = 1.0;
  S[i] = z;
  if (fout) {
    fout[i] = z.flags;
    __threadfence_system();
  }
}

// ===========================================================================
// host driver
// ===========================================================================
// ---- lane compaction for sparse masked operator launches -------------------
// A 64-lane group pays for a whole walk even when one lane computes.  When a
// masked A / A^T product selects only a few lanes (the 5-strike fallback
// usually does), their vectors are gathered out of the interleaved layout,
// multiplied by the W = 1 kernels and scattered back.  Every walk sum is an
// FP64 sum of vec_t values in both layouts, so the results are the same.
// Up to lane_limit() lanes (default 16, TV3_LANE_MAX=k sets another up to
// 64, 0 turns it off) go to the W = 1 kernels.  More lanes, when they fit in
// fewer 64-lane groups than the batch has, are packed into compact 64-lane
// groups (listed instance k in group k / 64, lane k % 64) and multiplied by
// the 64-lane kernels (TV3_COMPACT64=0 turns this off).  Lists hold at most
// kLaneCap instances.
constexpr int kLaneCap = 256;
struct LaneList {
  int n;
  int inst[kLaneCap];
};

// (read at every batch set-up)
int lane_limit() {
  const char *e = std::getenv("TV3_LANE_MAX");
  if (!e || !e[0]) return 16;
  return std::max(0, std::min(64, std::atoi(e)));
}

bool compact64_on() {
  const char *e = std::getenv("TV3_COMPACT64");
  return !(e && e[0] == '0');
}

// xc[k * pitch + e] = x[off + e * 64 + lane] for the k-th listed instance
__global__ __launch_bounds__(kBlock) void k_lane_gather(const GroupDesc *G, const vec_t *x,
                                                        vec_t *xc, LaneList L, int len,
                                                        int use_m, long long pitch) {
  const int k = blockIdx.y;
  if (k >= L.n) return;
  const int inst = L.inst[k];
  const GroupDesc &gd = G[inst / 64];
  const int b = inst % 64;
  const long long off = use_m ? gd.off_m : gd.off_n;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < len; e += gridDim.x * kBlock)
    xc[k * pitch + e] = x[off + (long long)e * 64 + b];
}

// y[off + e * 64 + lane] = yc[k * pitch + e] (- sub) for listed instances
// whose flags pass the launch's mask
__global__ __launch_bounds__(kBlock) void k_lane_scatter(const GroupDesc *G, const Scal *S,
                                                         const vec_t *yc, vec_t *y, LaneList L,
                                                         int len, int use_m, long long pitch,
                                                         uint32_t mask, const vec_t *sub) {
  const int k = blockIdx.y;
  if (k >= L.n) return;
  const int inst = L.inst[k];
  if (mask && (S[inst].flags & mask) != mask) return;
  const GroupDesc &gd = G[inst / 64];
  const int b = inst % 64;
  const long long off = use_m ? gd.off_m : gd.off_n;
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < len; e += gridDim.x * kBlock) {
    const long long o = off + (long long)e * 64 + b;
    const vec_t v = yc[k * pitch + e];
    y[o] = sub ? v - sub[o] : v;
  }
}

// compact 64-lane layout: xc[g * pitch + v * 64 + l] = x of listed instance
// 64 g + l at voxel (row) v; unlisted lanes of the last group get 0
__global__ __launch_bounds__(kBlock) void k_lane_gather64(const GroupDesc *G, const vec_t *x,
                                                          vec_t *xc, LaneList L, int len,
                                                          int use_m, long long pitch) {
  const int g = blockIdx.y;
  const long long E = (long long)len * 64;
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < E;
       e += (long long)gridDim.x * kBlock) {
    const int k = g * 64 + (int)(e & 63);
    vec_t v = 0.0f;
    if (k < L.n) {
      const int inst = L.inst[k];
      const GroupDesc &gd = G[inst / 64];
      v = x[(use_m ? gd.off_m : gd.off_n) + (e & ~63LL) + inst % 64];
    }
    xc[g * pitch + e] = v;
  }
}

// y of listed instances whose flags pass the mask = compact result (- sub)
__global__ __launch_bounds__(kBlock) void k_lane_scatter64(const GroupDesc *G, const Scal *S,
                                                           const vec_t *yc, vec_t *y, LaneList L,
                                                           int len, int use_m, long long pitch,
                                                           uint32_t mask, const vec_t *sub) {
  const int g = blockIdx.y;
  const long long E = (long long)len * 64;
  for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < E;
       e += (long long)gridDim.x * kBlock) {
    const int k = g * 64 + (int)(e & 63);
    if (k >= L.n) continue;
    const int inst = L.inst[k];
    if (mask && (S[inst].flags & mask) != mask) continue;
    const GroupDesc &gd = G[inst / 64];
    const long long o = (use_m ? gd.off_m : gd.off_n) + (e & ~63LL) + inst % 64;
    const vec_t v = yc[g * pitch + e];
    y[o] = sub ? v - sub[o] : v;
  }
}

struct Batch {
  int B = 0, G = 0, W = 1, Bpad = 0;
  std::vector<GroupDesc> groups;
  int max_n = 0, max_m = 0, kind = 0, dtype = 0;
  long long Nn = 0, Nm = 0;
  long long dblk_doubles = 0;  // slice partials of the blocked dense product
  bool z3 = false;
  bool layered = false;  // multi-layer instances: get_g takes two plain products
};

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// ---- autotuner for the dense W = 1 launch shape ------------------------------
// Each candidate is timed with HIP events on the batch's own operators (two
// launches after one warm-up); the fastest is cached per device and batch
// signature.  Candidates never change results (see DenseTune).
// TV3_AUTOTUNE=0 keeps the default shape.
using TuneKey = std::tuple<int, int, int, int, int, long long, int>;
std::mutex g_tune_mu;
std::map<TuneKey, DenseTune> g_tune_cache;

// blocks of a vector pass over the whole batch (TV3_VEC_BLOCKS overrides)
long long vec_blocks() {
  const char *e = std::getenv("TV3_VEC_BLOCKS");
  const long long v = e ? std::atoll(e) : 0;
  return v > 0 ? v : 2048;
}

// TV3_VEC4=0: 64-lane 2-D float batches run the vector passes with one
// instance per lane instead of four (k_descend4)
bool vec4_passes() {
  const char *e = std::getenv("TV3_VEC4");
  return !(e && e[0] == '0');
}

// TV3_SPIN=0: host-driven solves read the flags after a stream
// synchronization instead of polling the published launch numbers
bool spin_wait() {
  static const bool v = [] {
    const char *e = std::getenv("TV3_SPIN");
    return !(e && e[0] == '0');
  }();
  return v;
}

bool autotune_enabled() {
  const char *e = std::getenv("TV3_AUTOTUNE");
  return !(e && e[0] == '0');
}

bool layer_mfma_enabled() {
  const char *e = std::getenv("TV3_LAYER_MFMA");
  return !(e && e[0] == '0');
}

// TV3_DENSE_SHAPE=R,U,C forces the dense launch shape (R in {1,2,4} rows per
// wavefront of A x, U in {1,2,4,8} rows ahead, C in {4,8} columns per thread
// of the blocked A^T x), so that tests can run every candidate; false when
// unset or malformed
bool forced_dense_shape(DenseTune &t) {
  const char *e = std::getenv("TV3_DENSE_SHAPE");
  if (!e || !*e) return false;
  int r = 0, u = 0, c = 0;
  if (std::sscanf(e, "%d,%d,%d", &r, &u, &c) != 3) return false;
  if (!(r == 1 || r == 2 || r == 4) || !(u == 1 || u == 2 || u == 4 || u == 8) || !(c == 2 || c == 4 || c == 8))
    return false;
  t.rows = r;
  t.unroll = u;
  t.cols = c;
  return true;
}

// one kernel launch of the engine; with profiling on (and outside graph
// capture) an event pair around it adds its time to profile slot SLOT
#define TV3_GO(SLOT, ...)                      \
  do {                                         \
    if (!capturing_) prof_begin(ctx_, SLOT);   \
    hipLaunchKernelGGL(__VA_ARGS__);           \
    if (!capturing_) prof_end(ctx_);           \
  } while (0)

template <int W, class TS>
class Engine {
 public:
  Engine(tv3_ctx *ctx, const Batch &bt, const tv3_options &o)
      : ctx_(ctx), s_(ctx->stream), bt_(bt) {
    O_.mu = o.mu; O_.beta = o.beta; O_.mu0 = o.mu0; O_.beta0 = o.beta0;
    O_.tol = o.tol; O_.tol_inn = o.tol_inn; O_.maxit = o.maxit;
    O_.maxcnt = o.maxcnt; O_.gam = o.gam; O_.gamma = o.gamma;
    O_.rate_gam = o.rate_gam; O_.rate_cnt = o.rate_cnt; O_.c = o.c;
    O_.nonneg = o.nonneg; O_.tv_iso = o.tv_iso; O_.dt_exact = o.dt_exact;
    O_.numInitIterations = o.numInitIterations;
    O_.skipMulRatio = o.skipMulRatio;
    O_.maxRelativeChange = o.maxRelativeChange;
    O_.skip = (o.skipMulRatio != 0 && bt.kind != kDense) ? 1 : 0;
    O_.ls_exact = o.ls_exact;
    graph_ = o.graph != 0;
    isreal_ = o.isreal != 0;
    submat_mb_ = o.submat_mb > 0 ? o.submat_mb : 5;
    // blocks per group of the vector passes: about vec_blocks() over the
    // whole batch, at least 128 per group, at most kNblkGroup for 16/64-lane
    // groups (512 measured best on the 256 ring at B = 128 and 256: 1024 gave
    // 503 / 559 recons/s, 512 gave 511 / 579) and kNblkMax for one instance
    // (a 128^3 volume then runs 8 wavefronts per SIMD instead of 2)
    const long long cap = W == 1 ? kNblkMax : kNblkGroup;
    const long long per = std::max<long long>(128, vec_blocks() / std::max(bt.G, 1));
    nblk_n_ = (int)std::min<long long>(std::min<long long>(cap, per),
                                       std::max<long long>(1, ((long long)bt.max_n * W + kBlock - 1) / kBlock));
    nblk_m_ = (int)std::min<long long>(std::min<long long>(cap, per),
                                       std::max<long long>(1, ((long long)bt.max_m * W + kBlock - 1) / kBlock));
  }

  // dense W = 1 scratch: split-K partials of the plain A^T product or slice
  // partials of the blocked one (never live at the same time)
  size_t scratch_doubles() const {
    return (size_t)std::max<long long>((long long)bt_.G * (kCplx ? 32 : 16) * bt_.max_n, bt_.dblk_doubles);
  }

  int setup() {
    // 22 common n-vectors + 5 z-direction ones (3-D) + a work vector, rcdU, Uout
    const int nn_arrays = bt_.z3 ? 30 : 25;
    const int nm_arrays = 7;
    size_t need = 0;
    need += align_up(sizeof(vec_t) * bt_.Nn) * nn_arrays;
    need += align_up(sizeof(vec_t) * bt_.Nm) * nm_arrays;
    need += align_up(sizeof(double) * (size_t)bt_.Bpad * kNslot * kNblkMax);
    need += align_up(sizeof(MinMax) * (size_t)bt_.Bpad * kNblkMax);
    need += align_up(sizeof(Scal) * bt_.Bpad);
    need += align_up(sizeof(GroupDesc) * bt_.G);
    need += align_up(sizeof(vec_t *) * bt_.Bpad) * 2;
    if (bt_.kind == kDense && W == 1)
      need += align_up(sizeof(double) * scratch_doubles());
    lane_max_ = lane_limit();
    compact_ok_ = (W == 64 && bt_.kind == kGeometry);
    pitch_ = ((long long)std::max(bt_.max_n, bt_.max_m) + 63) / 64 * 64;
    gc64_max_ = (compact_ok_ && compact64_on()) ? std::min(bt_.G - 1, kLaneCap / 64) : 0;
    pitch64_ = (long long)std::max(bt_.max_n, bt_.max_m) * 64;
    if (compact_ok_) {
      need += align_up(sizeof(GroupDesc) * std::max(lane_max_, 1));
      need += align_up(sizeof(vec_t) * (size_t)lane_max_ * pitch_) * 2;
      need += align_up(sizeof(GroupDesc) * std::max(gc64_max_, 1));
      need += align_up(sizeof(vec_t) * (size_t)gc64_max_ * pitch64_) * 2;
    }
    if (ctx_->ws_bytes < need) {
      if (ctx_->ws) {
        (void)hipStreamSynchronize(s_);
        (void)hipFree(ctx_->ws);
        ctx_->ws = nullptr;
        ctx_->ws_bytes = 0;
      }
      int rc = hip_check(ctx_, hipMalloc(&ctx_->ws, need), "workspace hipMalloc");
      if (rc) return rc;
      ctx_->ws_bytes = need;
    }
    char *p = (char *)ctx_->ws;
    auto take = [&](size_t bytes) { char *r = p; p += align_up(bytes); return r; };
    const size_t fn = sizeof(vec_t) * bt_.Nn, fm = sizeof(vec_t) * bt_.Nm;
    vec_t **nptrs[] = {&a_.U, &a_.Up, &a_.Ux, &a_.Uxp, &a_.Uy, &a_.Uyp,
                       &a_.g, &a_.gp, &a_.g2, &a_.g2p, &a_.dU, &a_.dUx,
                       &a_.dUy, &a_.dg, &a_.dg2, &a_.d, &a_.DtsAtd, &a_.Wx,
                       &a_.Wy, &a_.sx, &a_.sy, &a_.Atb};
    for (vec_t **q : nptrs) *q = (vec_t *)take(fn);
    if (bt_.z3) {
      vec_t **zp[] = {&a_.Uz, &a_.Uzp, &a_.dUz, &a_.Wz, &a_.sz};
      for (vec_t **q : zp) *q = (vec_t *)take(fn);
    } else {
      a_.Uz = a_.Uzp = a_.dUz = a_.Wz = a_.sz = nullptr;
    }
    a_.tmpn = (vec_t *)take(fn);
    a_.rcdU = (vec_t *)take(fn);
    a_.Uout = (vec_t *)take(fn);
    vec_t **mptrs[] = {&a_.b, &a_.Au, &a_.Aup, &a_.dAu, &a_.delta, &a_.lastAu, &a_.tmpm};
    for (vec_t **q : mptrs) *q = (vec_t *)take(fm);
    part_ = (double *)take(sizeof(double) * (size_t)bt_.Bpad * kNslot * kNblkMax);
    mm_ = (MinMax *)take(sizeof(MinMax) * (size_t)bt_.Bpad * kNblkMax);
    S_ = (Scal *)take(sizeof(Scal) * bt_.Bpad);
    G_ = (GroupDesc *)take(sizeof(GroupDesc) * bt_.G);
    ptrs_in_ = (const vec_t **)take(sizeof(vec_t *) * bt_.Bpad);
    ptrs_out_ = (vec_t **)take(sizeof(vec_t *) * bt_.Bpad);
    scratch_ = (bt_.kind == kDense && W == 1) ? (vec_t *)take(sizeof(double) * scratch_doubles()) : nullptr;
    if (compact_ok_) {
      Gc_ = (GroupDesc *)take(sizeof(GroupDesc) * std::max(lane_max_, 1));
      xc_ = (vec_t *)take(sizeof(vec_t) * (size_t)lane_max_ * pitch_);
      yc_ = (vec_t *)take(sizeof(vec_t) * (size_t)lane_max_ * pitch_);
      gc_host_.assign(lane_max_, bt_.groups[0]);
      for (int k = 0; k < lane_max_; ++k) {
        gc_host_[k].off_n = k * pitch_;
        gc_host_[k].off_m = k * pitch_;
      }
      G64c_ = (GroupDesc *)take(sizeof(GroupDesc) * std::max(gc64_max_, 1));
      x64c_ = (vec_t *)take(sizeof(vec_t) * (size_t)gc64_max_ * pitch64_);
      y64c_ = (vec_t *)take(sizeof(vec_t) * (size_t)gc64_max_ * pitch64_);
      g64_host_.assign(gc64_max_, bt_.groups[0]);
      for (int k = 0; k < gc64_max_; ++k) {
        g64_host_[k].off_n = k * pitch64_;
        g64_host_[k].off_m = k * pitch64_;
      }
    }
    if (ctx_->hflags_n < (size_t)bt_.Bpad) {
      if (ctx_->hflags) {
        (void)hipStreamSynchronize(s_);
        (void)hipHostFree(ctx_->hflags);
        ctx_->hflags = ctx_->dflags = nullptr;
        ctx_->hflags_n = 0;
      }
      void *h = nullptr;
      int rc0 = hip_check(ctx_, hipHostMalloc(&h, sizeof(unsigned long long) * bt_.Bpad, hipHostMallocMapped | hipHostMallocCoherent), "flag hipHostMalloc");
      if (rc0) return rc0;
      void *d = nullptr;
      rc0 = hip_check(ctx_, hipHostGetDevicePointer(&d, h, 0), "flag device pointer");
      if (rc0) { (void)hipHostFree(h); return rc0; }
      std::memset(h, 0, sizeof(unsigned long long) * bt_.Bpad);
      ctx_->hflags = (unsigned long long *)h;
      ctx_->dflags = (unsigned long long *)d;
      ctx_->hflags_n = bt_.Bpad;
    }
    if ((size_t)(p - (char *)ctx_->ws) > need)
      return set_err(ctx_, TV3_ERR_ALLOC, "internal error: workspace layout exceeds its size");
    // zero all state (multipliers, W, DtsAtd start at 0; mat<T> zero-fills)
    int rc = hip_check(ctx_, hipMemsetAsync(ctx_->ws, 0, (size_t)(p - (char *)ctx_->ws), s_), "workspace clear");
    if (rc) return rc;
    rc = hip_check(ctx_, hipMemcpyAsync(G_, bt_.groups.data(), sizeof(GroupDesc) * bt_.G, hipMemcpyHostToDevice, s_), "group copy");
    if (!rc && compact_ok_ && lane_max_ > 0)
      rc = hip_check(ctx_, hipMemcpyAsync(Gc_, gc_host_.data(), sizeof(GroupDesc) * lane_max_, hipMemcpyHostToDevice, s_), "group copy");
    if (!rc && gc64_max_ > 0)
      rc = hip_check(ctx_, hipMemcpyAsync(G64c_, g64_host_.data(), sizeof(GroupDesc) * gc64_max_, hipMemcpyHostToDevice, s_), "group copy");
    return rc;
  }

  // times fn() (kTuneReps launches after a warm-up) on the solver stream, in ms
  static constexpr int kTuneReps = 8;
  template <class F>
  float time_launches(F fn) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 1e30f;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return 1e30f; }
    fn();
    (void)hipEventRecord(e0, s_);
    for (int k = 0; k < kTuneReps; ++k) fn();
    (void)hipEventRecord(e1, s_);
    float ms = 1e30f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) ms = 1e30f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
  }

  // picks the dense launch shape for this batch (dense, W = 1 only)
  int tune_dense() {
    // the FP64 kernels have one launch shape
    if (!(W == 1 && bt_.kind == kDense) || kOnlyW1) return TV3_OK;
    // Below kTuneMinBytes of stored A a product takes a few microseconds,
    // and event timings of such launches measure launch overhead more than
    // the kernel: fresh processes picked R = 1, 2 or 4 at random, and R = 4
    // ran 14 us inside the solve where R = 1 runs 5.5 us (64x64, m = 1229).
    // Small batches take the shape that was fastest inside the solve there.
    constexpr long long kTuneMinBytes = 256ll << 20;
    DenseTune best{1, 4, 8};
    // which blocked-A^T kernels the batch needs (per group: dblk_S > 0 takes
    // the row-slice pair, else the whole-row kernel); set with every choice
    bool any_slice = false, any_blk = false;
    for (const GroupDesc &g : bt_.groups) (g.dblk_S > 0 ? any_slice : any_blk) = true;
    // multi-layer batches: both products on the matrix cores
    // (TV3_LAYER_MFMA=0 keeps the per-layer W = 1 kernels)
    DenseTune mf{};
    if (bt_.layered && layer_mfma_enabled()) {
      mf.mfma = 1;
      mf.mfma_vec_f = mf.mfma_vec_b = 1;
      for (const GroupDesc &g : bt_.groups) {
        mf.mfma_layers = std::max(mf.mfma_layers, g.layers);
        mf.mfma_m = std::max(mf.mfma_m, g.plan.m);
        mf.mfma_n = std::max(mf.mfma_n, g.plan.n);
        if (!g.plan.a_vec) mf.mfma_vec_f = 0;
        if (g.plan.m % 4) mf.mfma_vec_b = 0;
      }
    }
    auto use = [&](DenseTune t) {
      t.no_blk = !any_blk;
      t.no_slice = !any_slice;
      t.mfma = mf.mfma;
      t.mfma_layers = mf.mfma_layers;
      t.mfma_m = mf.mfma_m;
      t.mfma_n = mf.mfma_n;
      t.mfma_vec_f = mf.mfma_vec_f;
      t.mfma_vec_b = mf.mfma_vec_b;
      dense_tune_set(t);
    };
    // the MFMA products have one launch shape
    if (mf.mfma) { use(best); return TV3_OK; }
    if (forced_dense_shape(best) || !autotune_enabled()) { use(best); return TV3_OK; }
    long long work = 0;
    for (const GroupDesc &g : bt_.groups) work += (long long)g.m * g.n;
    if (work * (bt_.dtype == TV3_BF16 ? 2 : 4) < kTuneMinBytes) { use(best); return TV3_OK; }
    const TuneKey key{ctx_->device, bt_.dtype, bt_.G, bt_.max_n, bt_.max_m, work, bt_.layered ? 1 : 0};
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune_cache.find(key);
      if (it != g_tune_cache.end()) { use(it->second); return TV3_OK; }
    }
    // every candidate is timed in kTuneRounds interleaved rounds and keeps
    // its fastest round
    constexpr int kTuneRounds = 3;
    const int rows_c[3] = {1, 2, 4};
    float tr[3] = {1e30f, 1e30f, 1e30f};
    for (int rd = 0; rd < kTuneRounds; ++rd)
      for (int i = 0; i < 3; ++i) {
        dense_tune_set(DenseTune{rows_c[i], best.unroll, best.cols});
        tr[i] = std::min(tr[i], time_launches([&] {
          launch_fwd<W>(s_, G_, bt_.G, bt_.max_m, bt_.kind, bt_.dtype, a_.Atb, a_.tmpm, flags_of(S_), 0);
        }));
      }
    best.rows = rows_c[std::min_element(tr, tr + 3) - tr];
    const int cols_c[3] = {8, 4, 2}, unr_c[4] = {1, 2, 4, 8};
    float tb[3][4];
    for (auto &row : tb)
      for (float &t : row) t = 1e30f;
    for (int rd = 0; rd < kTuneRounds; ++rd)
      for (int ci = 0; ci < 3; ++ci)
        for (int ui = 0; ui < 4; ++ui) {
          dense_tune_set(DenseTune{best.rows, unr_c[ui], cols_c[ci]});
          tb[ci][ui] = std::min(tb[ci][ui], time_launches([&] {
            launch_bwd<W>(s_, G_, bt_.G, bt_.max_n, bt_.kind, bt_.dtype, a_.tmpm, a_.tmpn, flags_of(S_), 0, nullptr, scratch_,
                          bt_.layered ? 0 : submat_mb_);
          }));
        }
    float tbest = 1e30f;
    for (int ci = 0; ci < 3; ++ci)
      for (int ui = 0; ui < 4; ++ui)
        if (tb[ci][ui] < tbest) { tbest = tb[ci][ui]; best.unroll = unr_c[ui]; best.cols = cols_c[ci]; }
    use(best);
    int rc = hip_check(ctx_, hipGetLastError(), "autotune launches");
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
    return TV3_OK;
  }

  int run(const vec_t *const *b, vec_t *const *U, const vec_t *const *Ut,
          tv3_info *infos, tv3_trace_rec *trace, int trace_cap) {
    // the MFMA layer products are this solve's choice only: later products
    // on this thread (tv3_gemv, single-layer solves) must not take them
    struct MfmaReset {
      ~MfmaReset() {
        DenseTune t = dense_tune_cur();
        t.mfma = 0;
        dense_tune_set(t);
      }
    } mfma_reset;
    int rc = setup();
    if (rc) return rc;
    if ((rc = tune_dense())) return rc;
    tv3_trace_rec *d_trace = nullptr;
    if (trace && trace_cap > 0) {
      rc = hip_check(ctx_, hipMalloc(&d_trace, sizeof(tv3_trace_rec) * (size_t)bt_.Bpad * trace_cap), "trace hipMalloc");
      if (rc) return rc;
      rc = hip_check(ctx_, hipMemsetAsync(d_trace, 0, sizeof(tv3_trace_rec) * (size_t)bt_.Bpad * trace_cap, s_), "trace clear");
      if (rc) { (void)hipFree(d_trace); return rc; }
    }
    trace_ = d_trace;
    trace_cap_ = trace_cap;
    rc = solve(b, U, Ut, infos);
    if (!rc && d_trace)
      rc = hip_check(ctx_, hipMemcpyAsync(trace, d_trace, sizeof(tv3_trace_rec) * (size_t)bt_.B * trace_cap, hipMemcpyDeviceToHost, s_), "trace copy");
    if (!rc) rc = hip_check(ctx_, hipStreamSynchronize(s_), "final sync");
    if (!rc && ctx_->prof) prof_collect(ctx_);
    if (d_trace) (void)hipFree(d_trace);
    return rc;
  }

 private:
  tv3_ctx *ctx_;
  hipStream_t s_;
  const Batch &bt_;
  OptDev O_{};
  int nblk_n_ = 1, nblk_m_ = 1;
  Arrs a_{};
  double *part_ = nullptr;
  MinMax *mm_ = nullptr;
  Scal *S_ = nullptr;
  GroupDesc *G_ = nullptr;
  const vec_t **ptrs_in_ = nullptr;
  vec_t **ptrs_out_ = nullptr;
  vec_t *scratch_ = nullptr;
  // graph mode (tv3_options::graph) and capture state
  bool graph_ = false, capturing_ = false;
  int submat_mb_ = 5;  // row blocks of the dense get_g product
  // lane compaction (W = 64, on-the-fly projector)
  bool compact_ok_ = false;
  int isreal_ = 1;  // complex data: keep the real part of the result
  int cmode_ = 0;  // lane compaction of the next products: 0 off, 1 W = 1, 2 compact W = 64
  int ncg_ = 0, gc64_max_ = 0;  // compact 64-lane groups in use / available
  long long pitch64_ = 0;
  GroupDesc *G64c_ = nullptr;
  vec_t *x64c_ = nullptr, *y64c_ = nullptr;
  std::vector<GroupDesc> g64_host_;
  int lane_max_ = 0;
  // the host copy of the flags is current: read_summary ran after the last
  // fin_* launch (true at the top of every step but the first)
  bool flags_fresh_ = false;
  LaneList lanes_{};
  GroupDesc *Gc_ = nullptr;
  vec_t *xc_ = nullptr, *yc_ = nullptr;
  long long pitch_ = 0;
  std::vector<GroupDesc> gc_host_;
  tv3_trace_rec *trace_ = nullptr;
  int trace_cap_ = 0;

  KA kn(uint32_t req, uint32_t forbid = 0) const { return KA{G_, S_, part_, nblk_n_, req, forbid}; }
  KA km(uint32_t req, uint32_t forbid = 0) const { return KA{G_, S_, part_, nblk_m_, req, forbid}; }
  dim3 gn() const { return dim3(nblk_n_, bt_.G); }
  dim3 gm() const { return dim3(nblk_m_, bt_.G); }
  // arguments of a fin_* kernel that publishes the flags (numbered: see
  // wait_published)
  FA fa(uint32_t req, uint32_t forbid = 0) const {
    FA f = fa_quiet(req, forbid);
    f.fout = ctx_->dflags;
    f.epoch = ++ctx_->pub_epoch;
    return f;
  }
  // fin_* kernels after which the host reads no flags need not publish them
  // to host memory
  FA fa_quiet(uint32_t req, uint32_t forbid = 0) const {
    FA f;
    f.S = S_; f.G = G_; f.part = part_; f.nblk_n = nblk_n_; f.nblk_m = nblk_m_;
    f.W = W; f.req = req; f.forbid = forbid; f.fout = nullptr; f.epoch = 0;
    f.trace = trace_; f.trace_cap = trace_cap_; f.O = O_;
    return f;
  }
  dim3 gi() const { return dim3(bt_.Bpad); }

  // algorithmic bytes of a profiled product launch: a host-driven solve of
  // one instance launches a product only while that instance needs it, so
  // the host adds its bytes and no counting kernel joins the timed stream.
  // F_BWD (the A^T-skip decision) is set on the device after the host's
  // last look, so those launches are counted on the device.
  void count_bytes(uint32_t mask, int trans) {
    if (!ctx_->prof) return;
    if (bt_.Bpad == 1 && !graph_ && !(mask & F_BWD)) prof_count_host(ctx_, bt_.groups[0], trans, (int)sizeof(vec_t));
    else prof_count(ctx_, G_, W, bt_.Bpad, flags_of(S_), mask, trans, (int)sizeof(vec_t));
  }

  void fwd(const vec_t *x, vec_t *y, uint32_t mask) {
    // lane compaction exists only in 64-lane engines (cmode_ stays 0 else)
    if constexpr (W == 64) {
      if (cmode_ == 2) {
        const dim3 gg(std::max(1, std::min(1024, (int)(((long long)bt_.max_n * 64 + kBlock - 1) / kBlock))), ncg_);
        const dim3 gs(std::max(1, std::min(1024, (int)(((long long)bt_.max_m * 64 + kBlock - 1) / kBlock))), ncg_);
        TV3_GO(kProfLanes, k_lane_gather64, gg, dim3(kBlock), 0, s_, G_, x, x64c_, lanes_, bt_.max_n, 0, pitch64_);
        prof_begin(ctx_, kProfGemvN);
        launch_fwd<64>(s_, G64c_, ncg_, bt_.max_m, bt_.kind, bt_.dtype, x64c_, y64c_, flags_of(S_), 0);
        prof_end(ctx_);
        prof_count_host(ctx_, bt_.groups[0], 0, (int)sizeof(vec_t), 64, ncg_, lanes_.n);
        TV3_GO(kProfLanes, k_lane_scatter64, gs, dim3(kBlock), 0, s_, G_, S_, y64c_, y, lanes_, bt_.max_m, 1, pitch64_, mask, (const vec_t *)nullptr);
        return;
      }
      if (cmode_ == 1) {
        const dim3 gg(std::max(1, std::min(64, (int)((bt_.max_n + kBlock - 1) / kBlock))), lanes_.n);
        const dim3 gs(std::max(1, std::min(64, (int)((bt_.max_m + kBlock - 1) / kBlock))), lanes_.n);
        TV3_GO(kProfLanes, k_lane_gather, gg, dim3(kBlock), 0, s_, G_, x, xc_, lanes_, bt_.max_n, 0, pitch_);
        // own slot: slots 10 / 11 (and their byte counts) keep the 64-lane
        // kernel's launches only
        prof_begin(ctx_, kProfGemvNLanes);
        launch_fwd<1>(s_, Gc_, lanes_.n, bt_.max_m, bt_.kind, bt_.dtype, xc_, yc_, flags_of(S_), 0);
        prof_end(ctx_);
        TV3_GO(kProfLanes, k_lane_scatter, gs, dim3(kBlock), 0, s_, G_, S_, yc_, y, lanes_, bt_.max_m, 1, pitch_, mask, (const vec_t *)nullptr);
        return;
      }
    }
    if (capturing_) {
      launch_fwd<W>(s_, G_, bt_.G, bt_.max_m, bt_.kind, bt_.dtype, x, y, flags_of(S_), mask);
      return;
    }
    prof_begin(ctx_, kProfGemvN);
    launch_fwd<W>(s_, G_, bt_.G, bt_.max_m, bt_.kind, bt_.dtype, x, y, flags_of(S_), mask);
    prof_end(ctx_);
    count_bytes(mask, 0);
  }
  // blocked: the get_g product (dense A: row blocks of submat_mb)
  void bwd(const vec_t *x, vec_t *y, uint32_t mask, const vec_t *sub, bool blocked = false) {
    // several layers: two plain products (2d/tval3_cpu_2d.cpp:365, 390-393)
    const int smb = (blocked && !bt_.layered) ? submat_mb_ : 0;
    if constexpr (W == 64) {
      if (cmode_ == 2) {
        const dim3 gg(std::max(1, std::min(1024, (int)(((long long)bt_.max_m * 64 + kBlock - 1) / kBlock))), ncg_);
        const dim3 gs(std::max(1, std::min(1024, (int)(((long long)bt_.max_n * 64 + kBlock - 1) / kBlock))), ncg_);
        TV3_GO(kProfLanes, k_lane_gather64, gg, dim3(kBlock), 0, s_, G_, x, x64c_, lanes_, bt_.max_m, 1, pitch64_);
        prof_begin(ctx_, kProfGemvT);
        launch_bwd<64>(s_, G64c_, ncg_, bt_.max_n, bt_.kind, bt_.dtype, x64c_, y64c_, flags_of(S_), 0, nullptr, nullptr, 0);
        prof_end(ctx_);
        prof_count_host(ctx_, bt_.groups[0], 1, (int)sizeof(vec_t), 64, ncg_, lanes_.n);
        TV3_GO(kProfLanes, k_lane_scatter64, gs, dim3(kBlock), 0, s_, G_, S_, y64c_, y, lanes_, bt_.max_n, 0, pitch64_, mask, sub);
        return;
      }
      if (cmode_ == 1) {
        const dim3 gg(std::max(1, std::min(64, (int)((bt_.max_m + kBlock - 1) / kBlock))), lanes_.n);
        const dim3 gs(std::max(1, std::min(64, (int)((bt_.max_n + kBlock - 1) / kBlock))), lanes_.n);
        TV3_GO(kProfLanes, k_lane_gather, gg, dim3(kBlock), 0, s_, G_, x, xc_, lanes_, bt_.max_m, 1, pitch_);
        prof_begin(ctx_, kProfGemvTLanes);
        launch_bwd<1>(s_, Gc_, lanes_.n, bt_.max_n, bt_.kind, bt_.dtype, xc_, yc_, flags_of(S_), 0, nullptr, nullptr, 0);
        prof_end(ctx_);
        TV3_GO(kProfLanes, k_lane_scatter, gs, dim3(kBlock), 0, s_, G_, S_, yc_, y, lanes_, bt_.max_n, 0, pitch_, mask, sub);
        return;
      }
    }
    if (capturing_) {
      launch_bwd<W>(s_, G_, bt_.G, bt_.max_n, bt_.kind, bt_.dtype, x, y, flags_of(S_), mask, sub, scratch_, smb);
      return;
    }
    prof_begin(ctx_, kProfGemvT);
    launch_bwd<W>(s_, G_, bt_.G, bt_.max_n, bt_.kind, bt_.dtype, x, y, flags_of(S_), mask, sub, scratch_, smb);
    prof_end(ctx_);
    count_bytes(mask, 1);
  }

  // instances whose published flags (last read_summary) contain all bits of
  // req, into lanes_, and the compaction mode for them: 1 (W = 1 kernels) up
  // to lane_max_ lanes, 2 (compact 64-lane groups) when they fill fewer
  // groups than the batch, else 0
  int pick_compact(uint32_t req) {
    const volatile unsigned long long *h = ctx_->hflags;
    lanes_.n = 0;
    for (int i = 0; i < bt_.Bpad; ++i) {
      if (((uint32_t)h[i] & req) != req) continue;
      if (lanes_.n == kLaneCap) return 0;
      lanes_.inst[lanes_.n++] = i;
    }
    if (lanes_.n == 0) return 0;
    if (lanes_.n <= lane_max_) return 1;
    ncg_ = (lanes_.n + 63) / 64;
    return ncg_ <= gc64_max_ ? 2 : 0;
  }

  // Waits until every instance's flags of the last publishing launch are in
  // host memory.  Host-driven solves poll the launch number the fin kernel
  // stores with the flags instead of synchronizing the stream: the host
  // wakes as soon as the flags land, and later launches need no drained
  // stream.  A captured graph replays its launch numbers, so graph mode (and
  // TV3_SPIN=0) synchronizes the stream; so does a poll that has not seen
  // the number after two seconds (a failed launch then reports its error).
  int wait_published() {
    if (graph_ || !spin_wait() || !ctx_->hflags)
      return hip_check(ctx_, hipStreamSynchronize(s_), "flag sync");
    const uint32_t e = ctx_->pub_epoch;
    const volatile unsigned long long *q = ctx_->hflags;
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < bt_.Bpad; ++i) {
      unsigned spins = 0;
      while ((uint32_t)(q[i] >> 32) != e) {
        if ((++spins & 1023) == 0 &&
            std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2))
          return hip_check(ctx_, hipStreamSynchronize(s_), "flag sync");
        __builtin_ia32_pause();
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return TV3_OK;
  }

  // OR of the instance flags the fin kernels published to mapped host memory
  int read_summary(uint32_t &out) {
    int rc = wait_published();
    out = 0;
    if (rc) return rc;
    flags_fresh_ = true;
    const volatile unsigned long long *h = ctx_->hflags;
    for (int i = 0; i < bt_.Bpad; ++i) out |= (uint32_t)h[i];
    return TV3_OK;
  }

  // get_g after U and Ux.. are in place: A U, A^T skip state, A^T, lam3/5
  void get_g_products(uint32_t req, uint32_t forbid, int force) {
    fwd(a_.U, a_.Au, req);
    {
      auto go = [&](auto kern) { TV3_GO(kProfGetG, kern, gm(), dim3(kBlock), 0, s_, km(req, forbid), a_); };
      const bool poly = !force && !O_.ls_exact;
      if (O_.skip)
        poly ? go(k_mpost<W, true, true>) : go(k_mpost<W, true, false>);
      else
        poly ? go(k_mpost<W, false, true>) : go(k_mpost<W, false, false>);
    }
    if (O_.skip) {
      TV3_GO(kProfScalars, fin_skip<TS>, gi(), dim3(kFin), 0, s_, fa_quiet(req, forbid), force);
      TV3_GO(kProfSetup, k_copy_m<W>, gm(), dim3(kBlock), 0, s_, km(req | F_BWD, forbid), a_.Au, a_.lastAu);
      bwd(a_.Au, a_.g, req | F_BWD, a_.Atb, true);
      TV3_GO(kProfSetup, k_skipg<W>, gn(), dim3(kBlock), 0, s_, kn(req | F_SKIPG, forbid), a_);
    } else {
      bwd(a_.Au, a_.g, req, a_.Atb, true);
    }
  }

  // descend + get_g's n-vector part, the instantiation for the batch's
  // dimensionality and line-search mode
  void descend(uint32_t req, bool poly) {
    auto go = [&](auto kern) {
      TV3_GO(kProfDescend, kern, gn(), dim3(kBlock), 0, s_, kn(req), a_, O_.nonneg, O_.dt_exact);
    };
    // four instances per lane for the step of every active instance (POLY);
    // the masked fallback launches, where most lanes skip, keep one instance
    // per lane (four per lane measured 348 -> 354 us per launch at B = 448,
    // against 386 -> 332 us for POLY at B = 512)
    if constexpr (W == 64 && std::is_same<vec_t, float>::value) {
      if (!bt_.z3 && poly && vec4_passes()) {
        TV3_GO(kProfDescend, (k_descend4<TS, true>), gn(), dim3(kBlock), 0, s_, kn(req), a_, O_.nonneg);
        return;
      }
    }
    if (bt_.z3)
      poly ? go(k_descend<W, TS, true, true>) : go(k_descend<W, TS, true, false>);
    else
      poly ? go(k_descend<W, TS, false, true>) : go(k_descend<W, TS, false, false>);
  }

  // update_W pass (INIT: the initial shrinkage, without the gradient)
  template <bool INIT>
  void wstep(uint32_t req) {
    auto go = [&](auto kern) {
      TV3_GO(kProfShrink, kern, gn(), dim3(kBlock), 0, s_, kn(req), a_, O_.dt_exact);
    };
    if constexpr (W == 64 && !INIT && std::is_same<vec_t, float>::value) {
      if (!bt_.z3 && vec4_passes()) {
        auto go4 = [&](auto kern) {
          TV3_GO(kProfShrink, kern, gn(), dim3(kBlock), 0, s_, kn(req), a_);
        };
        O_.tv_iso ? go4(k_wstep4<TS, true>) : go4(k_wstep4<TS, false>);
        return;
      }
    }
    if constexpr (kCplx) {  // anisotropic only (check_opts)
      bt_.z3 ? go(k_wstep<W, TS, INIT, true, false>) : go(k_wstep<W, TS, INIT, false, false>);
    } else {
      const bool iso = O_.tv_iso != 0;
      if (bt_.z3)
        iso ? go(k_wstep<W, TS, INIT, true, true>) : go(k_wstep<W, TS, INIT, true, false>);
      else
        iso ? go(k_wstep<W, TS, INIT, false, true>) : go(k_wstep<W, TS, INIT, false, false>);
    }
  }

  // tau of the first inner iteration / fallback: D(d), |d|^2, A d
  void tau_first(uint32_t req, uint32_t forbid) {
    if (bt_.z3)
      TV3_GO(kProfTau, (k_tau_first<W, true>), gn(), dim3(kBlock), 0, s_, kn(req, forbid), a_);
    else
      TV3_GO(kProfTau, (k_tau_first<W, false>), gn(), dim3(kBlock), 0, s_, kn(req, forbid), a_);
    fwd(a_.d, a_.tmpm, req);
    TV3_GO(kProfTau, k_msq<W>, gm(), dim3(kBlock), 0, s_, km(req, forbid), a_.tmpm);
  }

  // One inner iteration of every active instance (min_u, update_W, the stop
  // tests and, for instances whose inner loop ended, the outer update).  With
  // uniform = false the host skips the launches of branches no instance takes
  // (it reads the OR of the flags at each branch point); with uniform = true
  // every launch is issued, masked on the device by the same flags, and the
  // host reads nothing: the launch sequence is then the same every iteration
  // and is captured into a graph.
  int step(uint32_t &sum, bool uniform) {
    int rc = TV3_OK;
    const bool fresh = flags_fresh_;
    flags_fresh_ = false;
    // ---- min_u: get_tau (2d:533) ----
    if (uniform || (sum & F_FST)) {
      // instances restarting their inner loop: A d of a few lanes
      cmode_ = (!uniform && fresh && compact_ok_) ? pick_compact(F_ACTIVE | F_FST) : 0;
      tau_first(F_ACTIVE | F_FST, 0);
      cmode_ = 0;
    }
    TV3_GO(kProfScalars, fin_tau<TS>, gi(), dim3(kFin), 0, s_, fa_quiet(F_ACTIVE));
    // keep previous values (2d:536-537) by swapping buffers
    std::swap(a_.U, a_.Up); std::swap(a_.Ux, a_.Uxp); std::swap(a_.Uy, a_.Uyp);
    std::swap(a_.Uz, a_.Uzp); std::swap(a_.g, a_.gp); std::swap(a_.g2, a_.g2p);
    std::swap(a_.Au, a_.Aup);
    // descend (2d:540)
    descend(F_ACTIVE, !O_.ls_exact);
    get_g_products(F_ACTIVE, 0, 0);
    TV3_GO(kProfScalars, fin_descend<TS>, gi(), dim3(kFin), 0, s_, fa(F_ACTIVE), 0);
    if (!uniform && (rc = read_summary(sum))) return rc;
    // Armijo backtracking (2d:548-575)
    if (uniform || (sum & F_BT)) {
      if (!O_.ls_exact) {
        // step size already chosen on the device: rebuild the state once
        TV3_GO(kProfLineSearch, (k_interp_n<W, TS>), gn(), dim3(kBlock), 0, s_, kn(F_BT), a_);
        TV3_GO(kProfLineSearch, (k_interp_m<W, TS>), gm(), dim3(kBlock), 0, s_, km(F_BT), a_);
        TV3_GO(kProfScalars, fin_lams<TS>, gi(), dim3(kFin), 0, s_, fa_quiet(F_BT), 1);
      } else {
        TV3_GO(kProfLineSearch, k_bt_diff_n<W>, gn(), dim3(kBlock), 0, s_, kn(F_BT), a_);
        TV3_GO(kProfLineSearch, k_bt_diff_m<W>, gm(), dim3(kBlock), 0, s_, km(F_BT), a_);
        // at most five rounds: the fifth failure sets the fallback
        for (int round = 0; round < 5 && (uniform || (sum & F_BT)); ++round) {
          TV3_GO(kProfLineSearch, (k_bt_upd_n<W, TS>), gn(), dim3(kBlock), 0, s_, kn(F_BT), a_);
          TV3_GO(kProfLineSearch, (k_bt_upd_m<W, TS>), gm(), dim3(kBlock), 0, s_, km(F_BT), a_);
          TV3_GO(kProfScalars, fin_lams<TS>, gi(), dim3(kFin), 0, s_, fa(F_BT), 0);
          if (!uniform && (rc = read_summary(sum))) return rc;
        }
      }
    }
    if (uniform || (sum & F_FB)) {
      // 5-strike fallback (2d:561-567): tau from the first-iteration
      // formula, U = Up, descend(force)
      cmode_ = (!uniform && compact_ok_) ? pick_compact(F_FB) : 0;
      tau_first(F_FB, 0);
      TV3_GO(kProfScalars, fin_tau<TS>, gi(), dim3(kFin), 0, s_, fa_quiet(F_FB));
      descend(F_FB, false);
      get_g_products(F_FB, 0, 1);
      cmode_ = 0;
      TV3_GO(kProfScalars, fin_descend<TS>, gi(), dim3(kFin), 0, s_, fa_quiet(F_FB), 1);
    }
    // ---- update_W, reference values, gradient, stop test ----
    wstep<false>(F_ACTIVE);
    TV3_GO(kProfScalars, fin_wstep<TS>, gi(), dim3(kFin), 0, s_, fa(F_ACTIVE), 0);
    if (!uniform && (rc = read_summary(sum))) return rc;
    if (uniform || (sum & F_INNER_DONE)) {
      TV3_GO(kProfMlp, k_outer1<W>, gn(), dim3(kBlock), 0, s_, kn(F_INNER_DONE), a_);
      TV3_GO(kProfScalars, fin_outer1<TS>, gi(), dim3(kFin), 0, s_, fa_quiet(F_INNER_DONE));
      TV3_GO(kProfMlp, k_mlp_n<W>, gn(), dim3(kBlock), 0, s_, kn(F_MLP), a_);
      TV3_GO(kProfMlp, k_mlp_m<W>, gm(), dim3(kBlock), 0, s_, km(F_MLP), a_);
      TV3_GO(kProfScalars, fin_mlp<TS>, gi(), dim3(kFin), 0, s_, fa_quiet(F_MLP));
      TV3_GO(kProfMlp, k_grad<W>, gn(), dim3(kBlock), 0, s_, kn(F_MLP), a_, 1);
      TV3_GO(kProfSetup, (k_finish<W, TS>), gn(), dim3(kBlock), 0, s_, kn(F_JUST_DONE), a_, isreal_);
      TV3_GO(kProfScalars, fin_refresh, gi(), dim3(kFin), 0, s_, fa(0));
      if (!uniform && (rc = read_summary(sum))) return rc;
    }
    rc = hip_check(ctx_, hipGetLastError(), "step kernels");
    if (rc) return rc;
    return TV3_OK;
  }

  // Graph mode: two uniform iterations (the buffer swaps of one iteration
  // undo those of the other) captured once and replayed; the host reads the
  // flags after every `per_check` replays and stops when no instance is
  // active.  Replays after the last instance finished are fully masked.
  int run_graph(uint32_t &sum) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = hip_check(ctx_, hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    capturing_ = true;
    uint32_t dummy = ~0u;
    int rc1 = step(dummy, true);
    if (!rc1) rc1 = step(dummy, true);
    capturing_ = false;
    rc = hip_check(ctx_, hipStreamEndCapture(s_, &graph), "end capture");
    if (rc1 || rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc1 ? rc1 : rc;
    }
    rc = hip_check(ctx_, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    (void)hipGraphDestroy(graph);
    if (rc) return rc;
    const int per_check = bt_.W == 1 && bt_.B == 1 ? 4 : 1;
    while (!rc && (sum & F_ACTIVE)) {
      for (int k = 0; k < per_check && !rc; ++k)
        rc = hip_check(ctx_, hipGraphLaunch(exec, s_), "graph launch");
      if (!rc) rc = read_summary(sum);
    }
    (void)hipGraphExecDestroy(exec);
    return rc;
  }

  int solve(const vec_t *const *b, vec_t *const *U, const vec_t *const *Ut,
            tv3_info *infos) {
    int rc;
    const auto t0 = std::chrono::steady_clock::now();
    flags_fresh_ = false;
    // instance scalars and caller pointers
    TV3_GO(kProfScalars, k_init_scal, dim3((bt_.Bpad + 255) / 256), dim3(256), 0, s_, S_, bt_.B, bt_.Bpad, ctx_->dflags);
    std::vector<const vec_t *> hin(bt_.Bpad, nullptr);
    for (int i = 0; i < bt_.B; ++i) hin[i] = b[i];
    rc = hip_check(ctx_, hipMemcpyAsync(ptrs_in_, hin.data(), sizeof(vec_t *) * bt_.Bpad, hipMemcpyHostToDevice, s_), "ptr copy");
    if (rc) return rc;
    const dim3 gp_(std::max(1, std::min(nblk_m_, 1024)), bt_.G);
    TV3_GO(kProfSetup, k_pack<W>, gp_, dim3(kBlock), 0, s_, G_, ptrs_in_, a_.b, 1);
    // scaleb
    TV3_GO(kProfScaleb, k_absminmax<W>, gm(), dim3(kBlock), 0, s_, km(F_LIVE), a_.b, mm_);
    TV3_GO(kProfScaleb, fin_scaleb<TS>, gi(), dim3(kBlock), 0, s_, fa(F_LIVE), mm_, a_.b);
    TV3_GO(kProfScaleb, k_scale_m<W>, gm(), dim3(kBlock), 0, s_, km(F_LIVE), a_.b);
    // Atb = A^T bs; U = Atb
    bwd(a_.b, a_.Atb, F_LIVE, nullptr);
    TV3_GO(kProfSetup, k_init_U<W>, gn(), dim3(kBlock), 0, s_, kn(F_LIVE), a_);
    // D(U), shrinkage, get_g(force), get_gradient
    TV3_GO(kProfSetup, k_D<W>, gn(), dim3(kBlock), 0, s_, kn(F_LIVE), a_.U, a_.Ux, a_.Uy, a_.Uz);
    wstep<true>(F_LIVE);
    TV3_GO(kProfScalars, fin_wstep<TS>, gi(), dim3(kFin), 0, s_, fa(F_LIVE), 1);
    get_g_products(F_LIVE, 0, 1);
    TV3_GO(kProfScalars, fin_init2<TS>, gi(), dim3(kFin), 0, s_, fa(F_LIVE));
    TV3_GO(kProfMlp, k_grad<W>, gn(), dim3(kBlock), 0, s_, kn(F_LIVE), a_, 0);
    TV3_GO(kProfSetup, (k_finish<W, TS>), gn(), dim3(kBlock), 0, s_, kn(F_JUST_DONE), a_, isreal_);
    rc = hip_check(ctx_, hipGetLastError(), "init kernels");
    if (rc) return rc;

    uint32_t sum = (O_.maxcnt > 0 && O_.maxit > 0) ? (F_ACTIVE | F_FST) : 0u;
    if (!graph_) {
      while (sum & F_ACTIVE)
        if ((rc = step(sum, false))) return rc;
    } else if (sum & F_ACTIVE) {
      if ((rc = run_graph(sum))) return rc;
    }
    // outputs, rel_error
    std::vector<vec_t *> hout(bt_.Bpad, nullptr);
    for (int i = 0; i < bt_.B; ++i) hout[i] = U[i];
    rc = hip_check(ctx_, hipMemcpyAsync(ptrs_out_, hout.data(), sizeof(vec_t *) * bt_.Bpad, hipMemcpyHostToDevice, s_), "ptr copy");
    if (rc) return rc;
    const dim3 gpn(std::max(1, std::min(nblk_n_, 1024)), bt_.G);
    TV3_GO(kProfSetup, k_unpack<W>, gpn, dim3(kBlock), 0, s_, G_, a_.Uout, ptrs_out_);
    bool any_ut = false;
    std::vector<const vec_t *> hut(bt_.Bpad, nullptr);  // alive until the sync below
    if (Ut) {
      for (int i = 0; i < bt_.B; ++i) { hut[i] = Ut[i]; any_ut |= Ut[i] != nullptr; }
      if (any_ut) {
        rc = hip_check(ctx_, hipMemcpyAsync(ptrs_in_, hut.data(), sizeof(vec_t *) * bt_.Bpad, hipMemcpyHostToDevice, s_), "ptr copy");
        if (rc) return rc;
        TV3_GO(kProfSetup, k_pack<W>, gpn, dim3(kBlock), 0, s_, G_, ptrs_in_, a_.rcdU, 0);
        TV3_GO(kProfSetup, k_relerr<W>, gn(), dim3(kBlock), 0, s_, kn(F_LIVE), a_.rcdU, a_.Uout);
        TV3_GO(kProfScalars, fin_relerr<TS>, gi(), dim3(kFin), 0, s_, fa(F_LIVE));
      }
    }
    std::vector<Scal> hs(bt_.Bpad);
    rc = hip_check(ctx_, hipMemcpyAsync(hs.data(), S_, sizeof(Scal) * bt_.Bpad, hipMemcpyDeviceToHost, s_), "scalar copy");
    if (!rc) rc = hip_check(ctx_, hipStreamSynchronize(s_), "solve sync");
    if (rc) return rc;
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < bt_.B; ++i) {
      tv3_info &inf = infos[i];
      inf.secs = secs;
      inf.outer_iters = hs[i].count_outer;
      inf.total_iters = hs[i].count_total;
      inf.rel_chg = hs[i].RelChgOut;
      inf.rel_error = (Ut && Ut[i]) ? hs[i].rel_error : 0.0;
    }
    return TV3_OK;
  }
};

int check_opts(tv3_ctx *ctx, const tv3_options *o) {
  if (!o) return set_err(ctx, TV3_ERR_ARG, "Argument error: options missing");
  if (o->maxit < 0 || o->maxcnt < 0)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: maxit and maxcnt must be >= 0");
  if (kCplx && o->tv_iso)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: isotropic TV takes real data");
  return TV3_OK;
}

int solve_batch_impl(tv3_ctx *ctx, int B, const tv3_plan *const *A,
                     const tv3_options *o, const int *dims,
                     const vec_t *const *b, vec_t *const *U,
                     const vec_t *const *Ut, tv3_info *infos,
                     tv3_trace_rec *trace, int trace_cap) {
  if (!ctx) return TV3_ERR_ARG;
  if (B <= 0 || !A || !dims || !b || !U || !infos)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: batch arguments missing");
  int rc = check_opts(ctx, o);
  if (rc) return rc;
  if ((rc = hip_check(ctx, hipSetDevice(ctx->device), "hipSetDevice"))) return rc;
  Batch bt;
  bt.B = B;
  bool shared = true;
  std::vector<int> layers(B, 1);
  for (int i = 0; i < B; ++i) {
    if (!A[i] || !b[i] || !U[i]) return set_err(ctx, TV3_ERR_ARG, "Argument error: null plan or vector");
    const int q = dims[3 * i], p = dims[3 * i + 1], r = dims[3 * i + 2];
    if (q <= 0 || p <= 0 || r <= 0)
      return set_err(ctx, TV3_ERR_ARG, "Argument error: volume dimensions must be positive");
    // 2-D multi-layer mode (2d/tval3_cpu_2d.cpp:889-1016: layers = b.len /
    // A.dim_y): U holds a whole number of A.dim_x-pixel layers
    const long long nu = (long long)q * p * r, na = A[i]->dev.n;
    if (nu != na) {
      if (r != 1 || A[i]->dev.kind == kGeometry || na <= 0 || nu % na != 0)
        return set_err(ctx, TV3_ERR_ARG, "Argument error: the length of U must be equal to A.dim_x * layers!");
      if ((nu / na) * (long long)A[i]->dev.m > 0x7fffffffLL)
        return set_err(ctx, TV3_ERR_ARG, "Argument error: too many layers");
      layers[i] = (int)(nu / na);
      bt.layered = true;
    }
    if (!plan_ok(A[i]->dev)) return set_err(ctx, TV3_ERR_ARG, kPlanErr);
    if (A[i]->dev.kind != A[0]->dev.kind ||
        (A[i]->dev.kind == kDense && A[i]->dev.dtype != A[0]->dev.dtype))
      return set_err(ctx, TV3_ERR_ARG, "Argument error: one batch needs one operator kind and element type");
    if (i > 0 && (A[i] != A[0] || q != dims[0] || p != dims[1] || r != dims[2])) shared = false;
    if (r > 1) bt.z3 = true;
  }
  if (bt.layered && A[0]->dev.kind == kDense)
    for (int i = 0; i < B; ++i)
      if (layers[i] == 1)
        return set_err(ctx, TV3_ERR_ARG,
                       "Argument error: a dense batch cannot mix multi-layer and single-layer instances");
  bt.kind = A[0]->dev.kind;
  bt.dtype = A[0]->dev.dtype;
  // shared operator and size: 64-lane groups (one walk and one 256-byte
  // gather per voxel for 64 instances).  TV3_GROUP_WIDTH=16 selects 16-lane
  // groups dealt one per XCD (measured slower on the 256x256 ring).
  // Multi-layer instances run in W = 1 groups.
  // FP64 vectors: W = 1 groups only.  The 64-lane projector kernels address
  // vectors by 32-bit byte offsets of 256-byte rows: fewer than 2^24 voxels
  // and rays.
  const bool rows_ok = A[0]->dev.kind != kGeometry || (A[0]->dev.n < (1 << 24) && A[0]->dev.m < (1 << 24));
  bt.W = (shared && B > 1 && !bt.layered && !kOnlyW1 && rows_ok) ? 64 : 1;
  if (const char *ev = kOnlyW1 ? nullptr : getenv("TV3_GROUP_WIDTH")) {
    const int w = atoi(ev);
    if (shared && B > 1 && !bt.layered && rows_ok && (w == 16 || w == 64) && !(w == 16 && A[0]->dev.kind == kDense))
      bt.W = w;
  }
  const int W = bt.W;
  bt.G = (B + W - 1) / W;
  bt.Bpad = bt.G * W;
  long long on = 0, om = 0;
  for (int g = 0; g < bt.G; ++g) {
    const int i = g * W;
    GroupDesc gd{};
    gd.q = dims[3 * i]; gd.p = dims[3 * i + 1]; gd.r = dims[3 * i + 2];
    gd.layers = layers[i];
    gd.n = A[i]->dev.n * gd.layers; gd.m = A[i]->dev.m * gd.layers;
    gd.off_n = on; gd.off_m = om;
    gd.plan = A[i]->dev;
    // offsets stay multiples of 64 floats (256-byte aligned group vectors)
    on += ((long long)gd.n * W + 63) / 64 * 64;
    om += ((long long)gd.m * W + 63) / 64 * 64;
    bt.max_n = std::max(bt.max_n, gd.n);
    bt.max_m = std::max(bt.max_m, gd.m);
    if (W == 1 && gd.plan.kind == kDense && gd.layers == 1) {
      // blocked get_g product: groups with up to 64 row blocks split each
      // block into row slices (about 16 or more chunks in all)
      const int smb = o->submat_mb > 0 ? o->submat_mb : 5;
      gd.dblk_rows = dense_block_rows(smb, gd.n, (int)sizeof(vec_t));
      const long long nb = ((long long)gd.m + gd.dblk_rows - 1) / gd.dblk_rows;
      if (nb <= 64) {
        gd.dblk_S = (int)std::max<long long>(1, 16 / std::max<long long>(nb, 1));
        gd.dblk_off = bt.dblk_doubles;
        bt.dblk_doubles += nb * gd.dblk_S * (long long)gd.n;
      }
    }
    bt.groups.push_back(gd);
  }
  bt.Nn = std::max<long long>(on, 1);
  bt.Nm = std::max<long long>(om, 1);
#if TV3_SOLVER_VARIANT != 0
  // T_scal = the component type (the reference's double and complex
  // overloads); scal_double has no effect
  (void)W;
  Engine<1, vr_t> e(ctx, bt, *o);
  return e.run(b, U, Ut, infos, trace, trace_cap);
#else
  if (W == 64) {
    if (o->scal_double) {
      Engine<64, double> e(ctx, bt, *o);
      return e.run(b, U, Ut, infos, trace, trace_cap);
    }
    Engine<64, vec_t> e(ctx, bt, *o);
    return e.run(b, U, Ut, infos, trace, trace_cap);
  }
  if (W == 16) {
    if (o->scal_double) {
      Engine<16, double> e(ctx, bt, *o);
      return e.run(b, U, Ut, infos, trace, trace_cap);
    }
    Engine<16, vec_t> e(ctx, bt, *o);
    return e.run(b, U, Ut, infos, trace, trace_cap);
  }
  if (o->scal_double) {
    Engine<1, double> e(ctx, bt, *o);
    return e.run(b, U, Ut, infos, trace, trace_cap);
  }
  Engine<1, vec_t> e(ctx, bt, *o);
  return e.run(b, U, Ut, infos, trace, trace_cap);
#endif
}

#if TV3_SOLVER_VARIANT == 0
// ---------------------------------------------------------------------------
// standalone operators (single volume, W = 1)
// ---------------------------------------------------------------------------
template <class TS>
__global__ __launch_bounds__(kBlock) void k_shrink_only(KA K, Arrs a, int iso) {
  double acc[1] = {0};
  {
    constexpr int W = 1;
    NLOOP_BEGIN(1)
    const TS beta = (TS)S.beta, ib = (TS)1 / beta;
    const bool z3 = gd.r > 1;
    vec_t wx, wy, wz = 0.0f;
    TS t;
    if (iso) {
      t = shrink_iso<TS>(a.Ux[o], a.Uy[o], z3 ? a.Uz[o] : 0.0f, a.sx[o], a.sy[o],
                         z3 ? a.sz[o] : 0.0f, z3, beta, ib, wx, wy, wz);
    } else {
      TS tx, ty, tz = 0;
      wx = shrink1<TS>(a.Ux[o], a.sx[o], beta, ib, tx);
      wy = shrink1<TS>(a.Uy[o], a.sy[o], beta, ib, ty);
      if (z3) wz = shrink1<TS>(a.Uz[o], a.sz[o], beta, ib, tz);
      t = tx + ty;
      if (z3) t = t + tz;
    }
    a.Wx[o] = wx;
    a.Wy[o] = wy;
    if (z3) a.Wz[o] = wz;
    acc[0] += (double)t;
    (void)W;
    LOOP_END
  }
  const int sl[1] = {0};
  block_partials<1, 1>(acc, sl, K.part, 0, blockIdx.x, K.S, K.req, K.forbid);
}

__global__ __launch_bounds__(kBlock) void k_Dt_only(KA K, const vec_t *X,
                                                    const vec_t *Y,
                                                    const vec_t *Z, vec_t *res,
                                                    int exact) {
  constexpr int W = 1;
  NLOOP_BEGIN(1)
  const Vox w(v, gd.q, gd.p, gd.r);
  res[o] = dt_at(w, exact != 0, [&](int j) { return X[j]; },
                 [&](int j) { return Y[j]; }, [&](int j) { return Z[j]; });
  (void)W;
  LOOP_END
}

__global__ __launch_bounds__(kBlock) void k_dot_only(KA K, const vec_t *x,
                                                     const vec_t *y) {
  double acc[1] = {0};
  {
    constexpr int W = 1;
    NLOOP_BEGIN(1)
    acc[0] += dre(x[o], y[o]);
    (void)W;
    LOOP_END
  }
  const int sl[1] = {0};
  block_partials<1, 1>(acc, sl, K.part, 0, blockIdx.x, K.S, K.req, K.forbid);
}

// small single-volume scratch: one GroupDesc, one Scal, partials
struct OpScratch {
  GroupDesc *G;
  Scal *S;
  double *part;
  double *dense;
};

int op_scratch(tv3_ctx *ctx, const GroupDesc &gd, double beta, OpScratch &sc,
               long long dense_doubles) {
  const size_t need = align_up(sizeof(GroupDesc)) + align_up(sizeof(Scal)) +
                      align_up(sizeof(double) * kNslot * kNblkMax) +
                      align_up(sizeof(double) * std::max<long long>(dense_doubles, 1));
  if (ctx->ws_bytes < need) {
    if (ctx->ws) { (void)hipStreamSynchronize(ctx->stream); (void)hipFree(ctx->ws); ctx->ws = nullptr; ctx->ws_bytes = 0; }
    int rc = hip_check(ctx, hipMalloc(&ctx->ws, need), "workspace hipMalloc");
    if (rc) return rc;
    ctx->ws_bytes = need;
  }
  char *p = (char *)ctx->ws;
  sc.G = (GroupDesc *)p; p += align_up(sizeof(GroupDesc));
  sc.S = (Scal *)p; p += align_up(sizeof(Scal));
  sc.part = (double *)p; p += align_up(sizeof(double) * kNslot * kNblkMax);
  sc.dense = (double *)p;
  Scal s{};
  s.flags = F_LIVE | F_ACTIVE;
  s.beta = beta;
  // the sources are stack variables: wait for the copies to finish
  int rc = hip_check(ctx, hipMemcpyAsync(sc.G, &gd, sizeof(GroupDesc), hipMemcpyHostToDevice, ctx->stream), "copy");
  if (!rc) rc = hip_check(ctx, hipMemcpyAsync(sc.S, &s, sizeof(Scal), hipMemcpyHostToDevice, ctx->stream), "copy");
  if (!rc) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "copy sync");
  return rc;
}

int nblk_for(long long n) {
  return (int)std::min<long long>(kNblkGroup, std::max<long long>(1, (n + kBlock - 1) / kBlock));
}

int host_partial_sum(tv3_ctx *ctx, const double *part, int slot, int nblk, double *out) {
  std::vector<double> h(nblk);
  int rc = hip_check(ctx, hipMemcpyAsync(h.data(), part + (long long)slot * kNblkMax, sizeof(double) * nblk, hipMemcpyDeviceToHost, ctx->stream), "copy");
  if (!rc) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  if (rc) return rc;
  double s = 0;
  for (int i = 0; i < nblk; ++i) s += h[i];
  *out = s;
  return TV3_OK;
}

GroupDesc vol_desc(int q, int p, int r) {
  GroupDesc gd{};
  gd.q = q; gd.p = p; gd.r = r;
  gd.n = q * p * r;
  return gd;
}

// ---- the no-allocation gemv (tv3_gemv_noalloc) --------------------------------
// Its kernels take their group descriptor, instance record and, for dense
// A^T x, the kDenseChunks * n split-K partial sums from `scratch`, one array
// per GPU that lives as long as the process.  In host code `scratch` names a
// host-side stand-in: only the per-device address from hipGetSymbolAddress
// (g_noalloc_ptr) may reach a kernel or a copy.
constexpr int kNoAllocMaxN = 65536;                 // columns of a dense A^T x
constexpr size_t kNoAllocHead = 1024;               // GroupDesc, then the instance record
constexpr size_t kNoAllocBytes = kNoAllocHead + sizeof(double) * kDenseChunks * (size_t)kNoAllocMaxN;
static_assert(sizeof(GroupDesc) <= kNoAllocHead / 2 && sizeof(Scal) <= kNoAllocHead / 2, "no-alloc head layout");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoAllocDevices] = {};

// writes the call's descriptor and record (kernel arguments, so the call
// needs no host buffer, copy or stream synchronization)
__global__ void k_noalloc_head(GroupDesc gd, Scal s, GroupDesc *G, Scal *S) {
  if (threadIdx.x == 0) {
    *G = gd;
    *S = s;
  }
}
#endif  // TV3_SOLVER_VARIANT == 0

}  // namespace

#if TV3_SOLVER_VARIANT == 0
int tv3::noalloc_attach(int device) {
  if (device < 0 || device >= kNoAllocDevices) return TV3_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return TV3_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TV3_ERR_HIP;
  g_noalloc_ptr[device] = p;
  return TV3_OK;
}
#endif

// ===========================================================================
// C-ABI
// ===========================================================================
#if TV3_SOLVER_VARIANT != 0
#if TV3_SOLVER_VARIANT == 1
#define TV3_ENTRY(name) name##_f64
#elif TV3_SOLVER_VARIANT == 2
#define TV3_ENTRY(name) name##_c64
#else
#define TV3_ENTRY(name) name##_c128
#endif
namespace {
// caller vectors: vr_t components, vec_t elements (complex: re, im pairs)
bool vec_aligned(const void *p) { return p == nullptr || (uintptr_t)p % alignof(vec_t) == 0; }
}  // namespace

extern "C" {

int TV3_ENTRY(tv3_solve)(tv3_ctx *ctx, const tv3_plan *A, const tv3_options *opts, int dim_x,
                         int dim_y, int dim_z, const vr_t *b, vr_t *U, const vr_t *Ut,
                         tv3_info *info, tv3_trace_rec *trace, int trace_cap) {
  if (!ctx) return TV3_ERR_ARG;
  if (!A || !b || !U || !info) return set_err(ctx, TV3_ERR_ARG, "Argument error: null argument");
  if (!vec_aligned(b) || !vec_aligned(U) || !vec_aligned(Ut))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: vectors must be aligned to their element size");
  const int dims[3] = {dim_x, dim_y, dim_z};
  const tv3_plan *As[1] = {A};
  const vec_t *bs[1] = {(const vec_t *)b};
  vec_t *Us[1] = {(vec_t *)U};
  const vec_t *Uts[1] = {(const vec_t *)Ut};
  return solve_batch_impl(ctx, 1, As, opts, dims, bs, Us, Ut ? Uts : nullptr,
                          info, trace, trace_cap);
}

int TV3_ENTRY(tv3_solve_batch)(tv3_ctx *ctx, int B, const tv3_plan *const *A,
                               const tv3_options *opts, const int *dims,
                               const vr_t *const *b, vr_t *const *U,
                               const vr_t *const *Ut, tv3_info *infos) {
  if (!ctx) return TV3_ERR_ARG;
  if (B > 0 && b && U)
    for (int i = 0; i < B; ++i)
      if (!vec_aligned(b[i]) || !vec_aligned(U[i]) || (Ut && !vec_aligned(Ut[i])))
        return set_err(ctx, TV3_ERR_ARG, "Argument error: vectors must be aligned to their element size");
  return solve_batch_impl(ctx, B, A, opts, dims, (const vec_t *const *)b, (vec_t *const *)U,
                          (const vec_t *const *)Ut, infos, nullptr, 0);
}

int TV3_ENTRY(tv3_gemv)(tv3_ctx *ctx, const tv3_plan *A, int trans, const vr_t *x, vr_t *y) {
  if (!ctx) return TV3_ERR_ARG;
  if (!A || !x || !y) return set_err(ctx, TV3_ERR_ARG, "Argument error: null argument");
  if (!vec_aligned(x) || !vec_aligned(y))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: vectors must be aligned to their element size");
  if (!plan_ok(A->dev)) return set_err(ctx, TV3_ERR_ARG, kGemvErr);
  int rc = hip_check(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  GroupDesc gd{};
  gd.q = A->dev.n; gd.p = 1; gd.r = 1;
  gd.n = A->dev.n;
  gd.m = A->dev.m;
  gd.layers = 1;
  gd.plan = A->dev;
  // one group descriptor and, for dense A, the chunk sums of A^T x (two per
  // complex element)
  const long long dense = A->dev.kind == kDense ? (kCplx ? 32LL : 16LL) * A->dev.n : 1;
  const size_t need = align_up(sizeof(GroupDesc)) + align_up(sizeof(double) * dense);
  if (ctx->ws_bytes < need) {
    if (ctx->ws) { (void)hipStreamSynchronize(ctx->stream); (void)hipFree(ctx->ws); ctx->ws = nullptr; ctx->ws_bytes = 0; }
    if ((rc = hip_check(ctx, hipMalloc(&ctx->ws, need), "workspace hipMalloc"))) return rc;
    ctx->ws_bytes = need;
  }
  GroupDesc *G = (GroupDesc *)ctx->ws;
  vec_t *scratch = (vec_t *)((char *)ctx->ws + align_up(sizeof(GroupDesc)));
  rc = hip_check(ctx, hipMemcpyAsync(G, &gd, sizeof(GroupDesc), hipMemcpyHostToDevice, ctx->stream), "copy");
  // gd is a stack variable: wait for the copy
  if (!rc) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "copy sync");
  if (rc) return rc;
  const vec_t *xv = (const vec_t *)x;
  vec_t *yv = (vec_t *)y;
  if (trans)
    launch_bwd<1>(ctx->stream, G, 1, gd.n, A->dev.kind, A->dev.dtype, xv, yv, nullptr, 0, nullptr, scratch, 0);
  else
    launch_fwd<1>(ctx->stream, G, 1, gd.m, A->dev.kind, A->dev.dtype, xv, yv, nullptr, 0);
  return hip_check(ctx, hipGetLastError(), "gemv launch");
}

}  // extern "C"
#else
extern "C" {

int tv3_solve(tv3_ctx *ctx, const tv3_plan *A, const tv3_options *opts,
              int dim_x, int dim_y, int dim_z, const float *b, float *U,
              const float *Ut, tv3_info *info, tv3_trace_rec *trace,
              int trace_cap) {
  if (!ctx) return TV3_ERR_ARG;
  if (!A || !b || !U || !info) return set_err(ctx, TV3_ERR_ARG, "Argument error: null argument");
  const int dims[3] = {dim_x, dim_y, dim_z};
  const tv3_plan *As[1] = {A};
  const float *bs[1] = {b};
  float *Us[1] = {U};
  const float *Uts[1] = {Ut};
  return solve_batch_impl(ctx, 1, As, opts, dims, bs, Us, Ut ? Uts : nullptr,
                          info, trace, trace_cap);
}

int tv3_solve_batch(tv3_ctx *ctx, int B, const tv3_plan *const *A,
                    const tv3_options *opts, const int *dims,
                    const float *const *b, float *const *U,
                    const float *const *Ut, tv3_info *infos) {
  return solve_batch_impl(ctx, B, A, opts, dims, b, U, Ut, infos, nullptr, 0);
}

int tv3_gemv(tv3_ctx *ctx, const tv3_plan *A, int trans, const float *x,
             float *y) {
  if (!ctx) return TV3_ERR_ARG;
  if (!A || !x || !y) return set_err(ctx, TV3_ERR_ARG, "Argument error: null argument");
  if (!plan_takes_f32(A->dev))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: an f64 / double-valued or complex plan needs its own entry point (tv3_gemv_f64, tv3_gemv_c64, tv3_gemv_c128)");
  int rc = hip_check(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  GroupDesc gd = vol_desc(A->dev.n, 1, 1);
  gd.m = A->dev.m;
  gd.plan = A->dev;
  OpScratch sc;
  const long long dense = A->dev.kind == kDense ? 16LL * A->dev.n : 1;
  if ((rc = op_scratch(ctx, gd, 0.0, sc, dense))) return rc;
  if (trans)
    launch_bwd<1>(ctx->stream, sc.G, 1, gd.n, A->dev.kind, A->dev.dtype, x, y, sc.S, 0, nullptr, (float *)sc.dense, 0);
  else
    launch_fwd<1>(ctx->stream, sc.G, 1, gd.m, A->dev.kind, A->dev.dtype, x, y, sc.S, 0);
  return hip_check(ctx, hipGetLastError(), "gemv launch");
}

int tv3_gemv_noalloc(tv3_ctx *ctx, const tv3_plan *A, int trans, const float *x,
                     float *y) {
  if (!ctx) return TV3_ERR_ARG;
  if (!A || !x || !y) return set_err(ctx, TV3_ERR_ARG, "Argument error: null argument");
  if (!plan_takes_f32(A->dev))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: tv3_gemv_noalloc takes f32 / bf16 dense, float sparse or geometry plans");
  if (A->dev.kind == kDense && trans && A->dev.n > kNoAllocMaxN)
    return set_err(ctx, TV3_ERR_ARG,
                   "Argument error: tv3_gemv_noalloc: a dense A^T x needs n <= 65536 columns (its fixed scratch "
                   "holds 16 * 65536 FP64 partial sums); call tv3_gemv");
  if (ctx->device < 0 || ctx->device >= kNoAllocDevices || !g_noalloc_ptr[ctx->device])
    return set_err(ctx, TV3_ERR_HIP, "tv3_gemv_noalloc: the device scratch was not attached (tv3_ctx_create)");
  int rc = hip_check(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  char *base = (char *)g_noalloc_ptr[ctx->device];
  GroupDesc *G = (GroupDesc *)base;
  Scal *S = (Scal *)(base + kNoAllocHead / 2);
  double *part = (double *)(base + kNoAllocHead);
  GroupDesc gd = vol_desc(A->dev.n, 1, 1);
  gd.m = A->dev.m;
  gd.layers = 1;
  gd.plan = A->dev;
  Scal sc{};
  sc.flags = F_LIVE | F_ACTIVE;
  hipLaunchKernelGGL(k_noalloc_head, dim3(1), dim3(64), 0, ctx->stream, gd, sc, G, S);
  if (trans)
    launch_bwd<1>(ctx->stream, G, 1, gd.n, A->dev.kind, A->dev.dtype, x, y, flags_of(S), 0, nullptr, (float *)part, 0);
  else
    launch_fwd<1>(ctx->stream, G, 1, gd.m, A->dev.kind, A->dev.dtype, x, y, flags_of(S), 0);
  return hip_check(ctx, hipGetLastError(), "gemv launch");
}

int tv3_D(tv3_ctx *ctx, int q, int p, int r, const float *U, float *Ux,
          float *Uy, float *Uz) {
  if (!ctx) return TV3_ERR_ARG;
  if (q <= 0 || p <= 0 || r <= 0 || !U || !Ux || !Uy || (r > 1 && !Uz))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: D arguments");
  int rc = hip_check(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  OpScratch sc;
  const GroupDesc gd = vol_desc(q, p, r);
  if ((rc = op_scratch(ctx, gd, 0.0, sc, 1))) return rc;
  KA K{sc.G, sc.S, sc.part, nblk_for(gd.n), 0, 0};
  hipLaunchKernelGGL(k_D<1>, dim3(K.nblk, 1), dim3(kBlock), 0, ctx->stream, K, U, Ux, Uy, Uz);
  return hip_check(ctx, hipGetLastError(), "D launch");
}

int tv3_Dt(tv3_ctx *ctx, int q, int p, int r, int exact, const float *X,
           const float *Y, const float *Z, float *res) {
  if (!ctx) return TV3_ERR_ARG;
  if (q <= 0 || p <= 0 || r <= 0 || !X || !Y || !res || (r > 1 && !Z))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: Dt arguments");
  int rc = hip_check(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  OpScratch sc;
  const GroupDesc gd = vol_desc(q, p, r);
  if ((rc = op_scratch(ctx, gd, 0.0, sc, 1))) return rc;
  KA K{sc.G, sc.S, sc.part, nblk_for(gd.n), 0, 0};
  hipLaunchKernelGGL(k_Dt_only, dim3(K.nblk, 1), dim3(kBlock), 0, ctx->stream, K, X, Y, Z, res, exact);
  return hip_check(ctx, hipGetLastError(), "Dt launch");
}

int tv3_shrink(tv3_ctx *ctx, int q, int p, int r, double beta, int iso,
               int scal_double, const float *Ux, const float *Uy,
               const float *Uz, const float *sx, const float *sy,
               const float *sz, float *Wx, float *Wy, float *Wz,
               double *lam1) {
  if (!ctx) return TV3_ERR_ARG;
  if (q <= 0 || p <= 0 || r <= 0 || !Ux || !Uy || !sx || !sy || !Wx || !Wy ||
      (r > 1 && (!Uz || !sz || !Wz)) || !lam1)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: shrink arguments");
  int rc = hip_check(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  OpScratch sc;
  const GroupDesc gd = vol_desc(q, p, r);
  const double bt = scal_double ? beta : (double)(float)beta;
  if ((rc = op_scratch(ctx, gd, bt, sc, 1))) return rc;
  KA K{sc.G, sc.S, sc.part, nblk_for(gd.n), 0, 0};
  Arrs a{};
  a.Ux = (float *)Ux; a.Uy = (float *)Uy; a.Uz = (float *)Uz;
  a.sx = (float *)sx; a.sy = (float *)sy; a.sz = (float *)sz;
  a.Wx = Wx; a.Wy = Wy; a.Wz = Wz;
  if (scal_double)
    hipLaunchKernelGGL(k_shrink_only<double>, dim3(K.nblk, 1), dim3(kBlock), 0, ctx->stream, K, a, iso);
  else
    hipLaunchKernelGGL(k_shrink_only<float>, dim3(K.nblk, 1), dim3(kBlock), 0, ctx->stream, K, a, iso);
  if ((rc = hip_check(ctx, hipGetLastError(), "shrink launch"))) return rc;
  double s;
  if ((rc = host_partial_sum(ctx, sc.part, 0, K.nblk, &s))) return rc;
  *lam1 = scal_double ? s : (double)(float)s;
  return TV3_OK;
}

int tv3_dot(tv3_ctx *ctx, long long n, const float *x, const float *y,
            double *out) {
  if (!ctx) return TV3_ERR_ARG;
  if (n < 0 || n > 0x7fffffffLL || !x || !y || !out)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: dot arguments");
  int rc = hip_check(ctx, hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  if (n == 0) { *out = 0; return TV3_OK; }
  OpScratch sc;
  const GroupDesc gd = vol_desc((int)n, 1, 1);
  if ((rc = op_scratch(ctx, gd, 0.0, sc, 1))) return rc;
  KA K{sc.G, sc.S, sc.part, nblk_for(n), 0, 0};
  hipLaunchKernelGGL(k_dot_only, dim3(K.nblk, 1), dim3(kBlock), 0, ctx->stream, K, x, y);
  if ((rc = hip_check(ctx, hipGetLastError(), "dot launch"))) return rc;
  return host_partial_sum(ctx, sc.part, 0, K.nblk, out);
}

int tv3_nrm2(tv3_ctx *ctx, long long n, const float *x, double *out) {
  double s = 0;
  const int rc = tv3_dot(ctx, n, x, x, &s);
  if (!rc) *out = std::sqrt(s);
  return rc;
}

}  // extern "C"
#endif  // TV3_SOLVER_VARIANT
