// tv3_device.h — device helpers shared by the operator and solver kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "tv3_internal.h"

namespace tv3 {

// Load through a pointer that a kernel read from a descriptor (a plan's
// arrays).  The compiler cannot tell that such a pointer addresses global
// memory and emits a flat load, which also counts against the LDS counter:
// every later LDS access or cross-lane shuffle then waits for it.  Through
// the global address space it is an ordinary vector load.
template <class T>
__device__ __forceinline__ T ldg(const T *p) {
  return *(const __attribute__((address_space(1))) T *)p;
}
__device__ __forceinline__ RayDesc ldg(const RayDesc *p) {
  const int *q = (const int *)p;
  RayDesc r;
  r.lin0 = ldg(q);
  r.L = ldg(q + 1);
  r.s_major = ldg(q + 2);
  r.s_a = ldg(q + 3);
  r.s_b = ldg(q + 4);
  r.ma = ldg(q + 5);
  r.mb = ldg(q + 6);
  r.inv2L = __int_as_float(ldg(q + 7));
  return r;
}

// ---- Bresenham closed form --------------------------------------------------
// Number of minor-axis steps after j major steps of the reference's error
// recurrence (3d/mat_mult.h:241-359: error starts at 2*m - L, a minor step
// happens when it is > 0, then +2*m each step and -2*L per minor step):
//   k(j) = ceil((2*m*j - L) / (2*L)) = floor((2*m*j + L - 1) / (2*L)).
// The quotient is taken from a float reciprocal and fixed up by one step;
// plans keep 2*m*j + L below 2^24, where the estimate is within one.
__device__ __forceinline__ int bres_k(int m, int L, float inv2L, int j) {
  if (m == 0) return 0;
  const int num = 2 * m * j + L - 1;
  const int den = 2 * L;
  int q = (int)((float)num * inv2L);
  if (q * den > num)
    --q;
  else if ((q + 1) * den <= num)
    ++q;
  return q;
}

__device__ __forceinline__ int ray_lin(const RayDesc &r, int j) {
  return r.lin0 + j * r.s_major + bres_k(r.ma, r.L, r.inv2L, j) * r.s_a +
         bres_k(r.mb, r.L, r.inv2L, j) * r.s_b;
}

// ---- wavefront / block reductions (fixed order) ------------------------------
// One DPP lane move of a double (both 32-bit halves, same control); lanes
// whose source is outside the row, or whose row is not in RM, read 0.
template <int CTRL, int RM = 0xf>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, RM, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, RM, 0xf, true);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes, left in lane 63, by register moves (no LDS round
// trips): an inclusive scan of each 16-lane row (row_shr 1, 2, 4, 8), then
// the row totals (row_bcast 15 into rows 1 and 3, row_bcast 31 into rows 2
// and 3).  The tree is fixed, so the result does not depend on timing.
__device__ __forceinline__ double wave_total63(double v) {
  v += dpp_f64<0x111>(v);
  v += dpp_f64<0x112>(v);
  v += dpp_f64<0x114>(v);
  v += dpp_f64<0x118>(v);
  v += dpp_f64<0x142, 0xa>(v);
  v += dpp_f64<0x143, 0xc>(v);
  return v;
}

__device__ __forceinline__ double lane63(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// the wavefront sum, returned to every lane
__device__ __forceinline__ double wave_sum(double v) { return lane63(wave_total63(v)); }

// Sums NS per-thread values over the threads of one lane class (tid % W) of a
// 256-thread block and writes the block partial of instance inst0 + b to
// part[((inst0 + b) * kNslot + slot[s]) * kNblkMax + blk].  Order: inside
// each wavefront the DPP tree of wave_total63 (W = 1) or a butterfly over
// the lanes of one class, then wavefronts 0..3 in order.
__device__ __forceinline__ bool sel(uint32_t f, uint32_t req, uint32_t forbid) {
  return (f & req) == req && !(f & forbid);
}

// Only lanes whose instance passes the (req, forbid) flag test write, so a
// masked kernel never overwrites the partials another kernel left for the
// instances it skips.
//
// LDS holds one value per (wavefront, slot) for W = 1; for W = 64 the four
// wavefronts add into one array in turn (the same left-to-right order), so
// the kernels' occupancy is not bounded by this scratch (NS * 512 bytes
// instead of NS * 2 KiB).
template <int W, int NS, class SC>
__device__ __forceinline__ void block_partials(const double (&v)[NS],
                                               const int (&slot)[NS],
                                               double *part, int inst0,
                                               int blk, const SC *S,
                                               uint32_t req, uint32_t forbid) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if constexpr (W == 64) {
    __shared__ double acc[NS][64];
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
      if (wv == w) {
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[s][lane] = w == 0 ? v[s] : acc[s][lane] + v[s];
      }
      __syncthreads();
    }
    for (int t = threadIdx.x; t < NS * W; t += kBlock) {
      const int s = t / W, b = t % W;
      if (!sel(S[inst0 + b].flags, req, forbid)) continue;
      part[((long long)(inst0 + b) * kNslot + slot[s]) * kNblkMax + blk] = acc[s][b];
    }
  } else {
    __shared__ double red[4][NS][W];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double x = v[s];
      if constexpr (W == 1) {
        x = wave_total63(x);
        if (lane == 63) red[wv][s][0] = x;
      } else {
#pragma unroll
        for (int off = 32; off >= W; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane < W) red[wv][s][lane] = x;
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NS * W; t += kBlock) {
      const int s = t / W, b = t % W;
      if (!sel(S[inst0 + b].flags, req, forbid)) continue;
      const double x = ((red[0][s][b] + red[1][s][b]) + red[2][s][b]) + red[3][s][b];
      part[((long long)(inst0 + b) * kNslot + slot[s]) * kNblkMax + blk] = x;
    }
  }
}

// Sum of nblk partials of one slot, by a 256-thread block (fixed order);
// the result is returned to every thread.
__device__ __forceinline__ double block_sum_partials(const double *p, int nblk) {
  __shared__ double red[kBlock];
  double x = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kBlock) x += p[i];
  red[threadIdx.x] = x;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// ---- grid geometry of the reconstruction volume -------------------------------
// x is the contiguous axis; all differences are periodic (2d/tval3_cpu_2d.cpp:
// 268-278, 3d/tval3_cpu_3d.cpp:365-389).
struct Vox {
  int v, x, y, z;
  int q, p, r, s;
  __device__ __forceinline__ Vox(int v_, int q_, int p_, int r_)
      : v(v_), q(q_), p(p_), r(r_), s(q_ * p_) {
    x = v % q;
    const int t = v / q;
    y = t % p;
    z = t / p;
  }
  __device__ __forceinline__ int xn() const { return x < q - 1 ? v + 1 : v - (q - 1); }
  __device__ __forceinline__ int xp() const { return x > 0 ? v - 1 : v + (q - 1); }
  __device__ __forceinline__ int yn() const { return y < p - 1 ? v + q : v - (p - 1) * q; }
  __device__ __forceinline__ int yp() const { return y > 0 ? v - q : v + (p - 1) * q; }
  __device__ __forceinline__ int zn() const { return z < r - 1 ? v + s : v - (r - 1) * s; }
  __device__ __forceinline__ int zp() const { return z > 0 ? v - s : v + (r - 1) * s; }
};

// Dt at one voxel, given accessors VX(j), VY(j), VZ(j) of the three
// components.  The additions follow the reference's pass order
// (2d/tval3_cpu_2d.cpp:280-293, 3d/tval3_cpu_3d.cpp:394-425), one rounding
// each.  For dim_z >= 2 the reference's bulk y-pass covers every i >= dim_x,
// so the first row of each slice z >= 1 also receives +Y[i - dim_x] - Y[i]
// (the entry i - dim_x is the last row of the previous slice) before the
// per-slice wrap pass; exact = true leaves that term out.
template <class FX, class FY, class FZ>
__device__ __forceinline__ auto dt_at(const Vox &w, bool exact, FX VX, FY VY,
                                      FZ VZ) -> decltype(VX(0)) {
  decltype(VX(0)) t = VX(w.xp()) - VX(w.v);
  if (w.r == 1) {
    t = t + VY(w.yp());
    t = t - VY(w.v);
    return t;
  }
  if (w.v >= w.q && (w.y > 0 || !exact)) {
    t = t + VY(w.v - w.q);
    t = t - VY(w.v);
  }
  if (w.y == 0) {
    t = t + VY(w.v + (w.p - 1) * w.q);
    t = t - VY(w.v);
  }
  t = t + VZ(w.zp());
  t = t - VZ(w.v);
  return t;
}

// ---- shrinkage (2d/tval3_cpu_2d.cpp:588-611) ---------------------------------
// Anisotropic soft-threshold of one direction: ubar = u - sigma/beta rounded
// to the vector type (T_mat), w = max(|ubar| - 1/beta, 0) in T_scal,
// W = sign+(ubar) w.
template <class TS, class VC>
__device__ __forceinline__ VC shrink1(VC u, VC sg, TS beta, TS ib, TS &w) {
  const VC ub = (VC)((TS)u - (TS)sg / beta);
  TS t = fabs((TS)ub) - ib;
  t = (t >= (TS)0) ? t : (TS)0;
  w = t;
  return (VC)((ub >= (VC)0) ? t : -t);
}

// Isotropic group soft-threshold (no reference counterpart): ubar as above per
// direction, r = |ubar|_2, w = max(r - 1/beta, 0), W_dir = w * ubar_dir / r.
template <class TS, class VC>
__device__ __forceinline__ TS shrink_iso(VC ux, VC uy, VC uz, VC sx, VC sy, VC sz, bool z3,
                                         TS beta, TS ib, VC &Wx, VC &Wy, VC &Wz) {
  const VC bx = (VC)((TS)ux - (TS)sx / beta);
  const VC by = (VC)((TS)uy - (TS)sy / beta);
  const VC bz = z3 ? (VC)((TS)uz - (TS)sz / beta) : (VC)0;
  TS ss = (TS)bx * (TS)bx + (TS)by * (TS)by;
  if (z3) ss = ss + (TS)bz * (TS)bz;
  const TS rr = sqrt(ss);
  TS w = rr - ib;
  w = (w >= (TS)0) ? w : (TS)0;
  Wx = (rr > 0) ? (VC)(w * (TS)bx / rr) : (VC)0;
  Wy = (rr > 0) ? (VC)(w * (TS)by / rr) : (VC)0;
  Wz = (z3 && rr > 0) ? (VC)(w * (TS)bz / rr) : (VC)0;
  return w;
}

// Complex anisotropic shrinkage of one direction (2d/tval3_cpu_2d.cpp:614-649,
// 3d/tval3_cpu_3d.cpp:876-947): ubar = u - sigma/beta per component, rounded
// to the component type R; |ubar| = sqrt(re*re + im*im) in R; w = max(|ubar|
// - 1/beta, 0) in R.  cshrink_w gives W = w * b / |ubar| per component in R
// (0 where |ubar| = 0); the reference's 3-D Wz takes b from the y direction.
template <class C>
struct CShr {
  C bar;
  decltype(C::x) mag, w;
};
template <class TS, class C>
__device__ __forceinline__ CShr<C> cshrink_bar(C u, C sg, TS beta, TS ib) {
  typedef decltype(C::x) R;
  CShr<C> r;
  r.bar = C((R)((TS)u.x - (TS)sg.x / beta), (R)((TS)u.y - (TS)sg.y / beta));
  r.mag = sqrt(r.bar.x * r.bar.x + r.bar.y * r.bar.y);
  const R t = (R)((TS)r.mag - ib);
  r.w = (t >= (R)0) ? t : (R)0;
  return r;
}
template <class C, class R>
__device__ __forceinline__ C cshrink_w(R w, C b, R mag) {
  return (mag > (R)0) ? C(w * b.x / mag, w * b.y / mag) : C((R)0, (R)0);
}

}  // namespace tv3
