// tv3_plan.hip — context, measurement-operator plans and the A / A^T kernels.
//
// Reference counterparts: gemv() for mat<T> (2d/mat_mult.h:15-36),
// sparse_mat<T> (2d/mat_mult.h:38-91, 3d/mat_mult.h:47-118) and geometry
// (3d/mat_mult.h:148-595).  Every product accumulates in FP64 and rounds once
// to float; no kernel uses floating-point atomics, so results are
// deterministic run to run.  A^T of the on-the-fly projector gathers over a
// voxel-major transpose built once per plan (rays of each voxel in ascending
// order), which replaces the reference's per-thread private output buffers
// and serial reduction (3d/mat_mult.h:367-385, 573-581).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "tv3_internal.h"
#include "tv3_device.h"

using namespace tv3;

namespace tv3 {

int set_err(tv3_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

const char *prof_slot_name(int slot) {
  switch (slot) {
    case kProfScaleb: return "scaleb (k_absminmax, fin_scaleb, k_scale_m)";
    case kProfShrink: return "shrinkage with update_W, Dt, get_gradient and the BB dots (k_wstep)";
    case kProfGemvN: return "gemv CblasNoTrans: A x";
    case kProfGemvT: return "gemv CblasConjTrans: A^T x";
    case kProfDescend: return "descend with nonneg, D, get_g's Dt, lam2, lam4 (k_descend)";
    case kProfLineSearch: return "min_u backtracking, update_g (k_interp_*, k_bt_*)";
    case kProfTau: return "get_tau first step: D(d), |d|^2, |Ad|^2 (k_tau_first, k_msq)";
    case kProfGetG: return "get_g m-part: Au - b, lam3, lam5, skip norms (k_mpost)";
    case kProfMlp: return "update_mlp, outer step, get_gradient (k_mlp_*, k_outer1, k_grad)";
    case kProfScalars: return "per-instance scalar control (fin_*)";
    case kProfLanes: return "lane compaction of masked products (k_lane_*)";
    case kProfSetup: return "setup, copies and finish (k_pack, k_init_U, k_D, k_finish, ...)";
    case kProfGemvNLanes: return "gemv CblasNoTrans on compacted lanes (W = 1 kernels): A x";
    case kProfGemvTLanes: return "gemv CblasConjTrans on compacted lanes (W = 1 kernels): A^T x";
    default: return "";
  }
}

void prof_begin(tv3_ctx *ctx, int kind) {
  ctx->prof_open = false;
  if (!ctx->prof || kind < 0 || kind >= kProfSlots) return;
  if (ctx->prof_level != 2 && kind != kProfGemvN && kind != kProfGemvT) return;
  if (ctx->evkind.size() >= 4096) {
    // fold the finished pairs in, so that long solves reuse the pool
    hipStreamCaptureStatus cs = hipStreamCaptureStatusActive;
    if (hipStreamIsCapturing(ctx->stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone &&
        hipStreamSynchronize(ctx->stream) == hipSuccess)
      prof_collect(ctx);
  }
  const size_t i = ctx->evkind.size() * 2;
  while (ctx->evpool.size() < i + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { ctx->prof = false; return; }
    ctx->evpool.push_back(e);
  }
  ctx->evkind.push_back(kind);
  (void)hipEventRecord(ctx->evpool[i], ctx->stream);
  ctx->prof_open = true;
}

void prof_end(tv3_ctx *ctx) {
  if (!ctx->prof || !ctx->prof_open || ctx->evkind.empty()) return;
  (void)hipEventRecord(ctx->evpool[ctx->evkind.size() * 2 - 1], ctx->stream);
  ctx->prof_open = false;
}

// folds the recorded pairs into the per-kind totals (after a stream sync)
void prof_collect(tv3_ctx *ctx) {
  for (size_t k = 0; k < ctx->evkind.size(); ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->evpool[2 * k], ctx->evpool[2 * k + 1]) == hipSuccess) {
      ctx->prof_ms[ctx->evkind[k]] += ms;
      ctx->prof_n[ctx->evkind[k]] += 1;
    }
  }
  ctx->evkind.clear();
}

// Compulsory bytes of a product launch: the operator's stored or walked
// description once for every group that computes (dense A; CSR or CSC
// indices, values and row pointers; the projector's ray descriptors and
// walk order, or its voxel-major transpose lists), and each computing
// lane's input and output vectors (vbytes per element, every layer).
__host__ __device__ inline unsigned long long prof_op_bytes(const PlanDev &pd, int trans, int W) {
  const unsigned long long m = pd.m, n = pd.n, nnz = pd.nnz;
  if (pd.kind == kDense) {
    const unsigned long long es = pd.dtype == TV3_BF16 ? 2ull
                                  : pd.dtype == TV3_C128 ? 16ull
                                  : (pd.dtype == TV3_F64 || pd.dtype == TV3_C64) ? 8ull : 4ull;
    return es * m * n;
  }
  if (pd.kind == kSparse) {
    const unsigned long long vs = pd.cval == 3 ? 16ull : (pd.cval == 2 || pd.r_vald) ? 8ull : 4ull;
    return (4 + vs) * nnz + 4 * ((trans ? n : m) + 1);
  }
  // geometry: 32-byte RayDesc per ray (and the 64-lane kernel's ray order),
  // or t_ptr, t_col and t_ray of the transpose; the staged single-instance
  // transpose reads s_col, s_ptr, 2-byte s_loc entries (4-byte g_cray
  // entries in tiles that stage nothing), the tiles' ray lists and their
  // pointers instead
  if (trans && W == 1 && pd.s_col) {
    const unsigned long long tiles = (pd.s_nz + pd.s_V - 1) / pd.s_V;
    return 2ull * nnz + 2ull * (unsigned long long)pd.s_dnnz + 4ull * n + 4ull * (pd.s_nz + 1) +
           8ull * pd.s_nh + 4ull * (tiles + 1) + 4ull * (unsigned long long)pd.s_rays;
  }
  return trans ? 4ull * nnz + 8ull * n + 4ull : 32ull * m + (W == 64 ? 4ull * m : 0ull);
}
__host__ __device__ inline unsigned long long prof_vec_bytes(const GroupDesc &gd, int vbytes) {
  return (unsigned long long)vbytes * ((unsigned long long)gd.n + (unsigned long long)gd.m);
}

// one thread per group: acc[trans] += compulsory bytes, acc[2 + trans] +=
// lanes that computed
__global__ void k_prof_bytes(const GroupDesc *G, int W, int nG, const InstScal *S,
                             uint32_t mask, int trans, int vbytes, unsigned long long *acc) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nG) return;
  unsigned long long cnt = 0;
  for (int l = 0; l < W; ++l) {
    const uint32_t f = S[g * W + l].flags;
    if ((f & F_LIVE) && (!mask || (f & mask) == mask)) ++cnt;
  }
  if (!cnt) return;
  const GroupDesc &gd = G[g];
  atomicAdd(&acc[trans], prof_op_bytes(gd.plan, trans, W) + cnt * prof_vec_bytes(gd, vbytes));
  atomicAdd(&acc[2 + trans], cnt);
}

void prof_count_host(tv3_ctx *ctx, const GroupDesc &g, int trans, int vbytes, int W, int groups,
                     int lanes) {
  if (!ctx->prof) return;
  ctx->prof_hbytes[trans] += (unsigned long long)groups * prof_op_bytes(g.plan, trans, W) +
                             (unsigned long long)lanes * prof_vec_bytes(g, vbytes);
  ctx->prof_hbytes[2 + trans] += (unsigned long long)lanes;
}

void prof_count(tv3_ctx *ctx, const GroupDesc *G_dev, int W, int Bpad,
                const InstScal *S, uint32_t mask, int trans, int vbytes) {
  if (!ctx->prof || !ctx->prof_bytes || ctx->prof_level == 3) return;
  const int nG = Bpad / W;
  hipLaunchKernelGGL(k_prof_bytes, dim3((nG + 255) / 256), dim3(256), 0, ctx->stream,
                     G_dev, W, nG, S, mask, trans, vbytes, ctx->prof_bytes);
}

int hip_check(tv3_ctx *ctx, hipError_t e, const char *what) {
  if (e == hipSuccess) return TV3_OK;
  return set_err(ctx, e == hipErrorOutOfMemory ? TV3_ERR_ALLOC : TV3_ERR_HIP,
                 std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// forward product: y[row] = sum_k A[row, k] x[k]
// ---------------------------------------------------------------------------

// W = 1: L lanes per row (64 / L rows per wavefront); lane k of a row takes
// entries k, k+L, ... in order; the L lane sums are combined by a fixed
// butterfly.  Geometry and CSR rows issue kW1Unroll index and gather loads
// before they add.  kSubF / kSubT: lanes per row of A x / A^T y (8 and 8
// measured best on the 128^3 hemisphere and the 256 ring's plain path;
// four loads ahead took A^T y there from 0.250 to 0.231 ms; with the
// uncrossed voxels split off, A^T y of the hemisphere takes 0.186 ms at
// 8 lanes x 4 loads, 0.187 at 16 x 4, 0.193 at 16 x 2, 0.203 at 8 x 2,
// 0.235 at 4 x 4 and 0.251 at 32 x 2).
constexpr int kSubF = 8;
constexpr int kSubT = 8;
// geometry columns of A^T y in the plan's tiled voxel order (t_*): 0.224 ->
// 0.196 ms on the 128^3 hemisphere; the tiled ray order made A x slower
// there (0.091 -> 0.098 ms), so A x keeps the measurement order
constexpr bool kW1TiledF = false;
constexpr bool kW1TiledT = true;
constexpr int kW1Unroll = 4;

template <int L>
__device__ __forceinline__ double lane_sum(double v) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int L>
__global__ __launch_bounds__(kBlock) void k_fwd_w1(const GroupDesc *G,
                                                   const float *x, float *y,
                                                   const InstScal *S,
                                                   uint32_t mask) {
  const GroupDesc &gd = G[blockIdx.y];
  const int sl = threadIdx.x & (L - 1);
  const int row = blockIdx.x * (kBlock / L) + (threadIdx.x / L);
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const PlanDev &P = gd.plan;
  // row of A and layer of this output row (one layer: the row itself; a
  // geometry row slot walks the plan's tiled ray order)
  int ar = row % P.m;
  if (kW1TiledF && P.kind == kGeometry && P.ray_order && row < gd.m) ar = ldg(P.ray_order + row);
  const float *xg = x + gd.off_n + (long long)(row / P.m) * P.n;
  double acc = 0.0;
  if (row < gd.m) {
    constexpr int U = kW1Unroll;
    if (P.kind == kGeometry) {
      const RayDesc rd = ldg(P.rays + ar);
      const float sc = P.scale;
      for (int j = sl; j <= rd.L; j += L * U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int jj = j + u * L;
          v[u] = jj <= rd.L ? xg[ray_lin(rd, jj)] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (j + u * L <= rd.L) acc += (double)(v[u] * sc);
      }
    } else if (P.kind == kSparse) {
      const int k0 = ldg(P.r_ptr + ar), k1 = ldg(P.r_ptr + ar + 1);
      for (int k = k0 + sl; k < k1; k += L * U) {
        int id[U];
        float w[U], v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int kk = k + u * L;
          id[u] = kk < k1 ? ldg(P.r_ind + kk) : -1;
          w[u] = kk < k1 ? ldg(P.r_val + kk) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = id[u] >= 0 ? xg[id[u]] : 0.0f;
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (id[u] >= 0) acc += (double)w[u] * (double)v[u];
      }
    } else {
      const long long base = (long long)ar * P.n;
      if (P.dtype == TV3_F32) {
        const float *A = (const float *)P.A + base;
        for (int j = sl; j < P.n; j += L) acc += (double)A[j] * (double)xg[j];
      } else {
        const __hip_bfloat16 *A = (const __hip_bfloat16 *)P.A + base;
        for (int j = sl; j < P.n; j += L)
          acc += (double)__bfloat162float(A[j]) * (double)xg[j];
      }
    }
  }
  acc = lane_sum<L>(acc);
  if (sl == 0 && row < gd.m) y[gd.off_m + (long long)(row / P.m) * P.m + ar] = (float)acc;
}

// W = 64: one wavefront per row, one lane per instance.  The walk, the
// indices and the matrix entries are wavefront-uniform and are kept in scalar
// registers (readfirstlane / readlane), so the scalar unit runs the walk and
// the address arithmetic; every gather is one 256-byte coalesced load of the
// 64 instances' values of a voxel (scalar row base + lane offset).  Entries
// are consumed in order (the FP64 sum is sequential per lane) but loaded
// kPipe at a time so that kPipe gathers are in flight per wavefront.
constexpr int kPipe = 16;

__device__ __forceinline__ int uni_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

__global__ __launch_bounds__(kBlock) void k_fwd_w64(const GroupDesc *G,
                                                    const float *x, float *y,
                                                    const InstScal *S,
                                                    uint32_t mask) {
  const GroupDesc &gd = G[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int row = uni_i(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
  if (row >= gd.m) return;
  const bool on = !mask || (S[blockIdx.y * 64 + lane].flags & mask) == mask;
  if (!__any(on)) return;
  const PlanDev &P = gd.plan;
  const float *xb = x + gd.off_n;
  double acc = 0.0;
  if (P.kind == kGeometry) {
    // voxel indices of 64 consecutive walk steps, one per lane (closed form of
    // the reference's recurrence), handed to the whole wavefront by readlane
    const RayDesc rd = P.rays[row];
    const int L = uni_i(rd.L);
    const float sc = P.scale;
    const bool unit = sc == 1.0f;
    for (int j0 = 0; j0 <= L; j0 += 64) {
      const int jl = j0 + lane;
      const int myidx = jl <= L ? ray_lin(rd, jl) : 0;
      const int cnt = min(64, L + 1 - j0);
      for (int u0 = 0; u0 < cnt; u0 += kPipe) {
        float v[kPipe];
#pragma unroll
        for (int u = 0; u < kPipe; ++u) {
          const int iu = __builtin_amdgcn_readlane(myidx, u0 + u);
          v[u] = (u0 + u < cnt) ? (xb + (long long)iu * 64)[lane] : 0.0f;
        }
        if (unit) {
#pragma unroll
          for (int u = 0; u < kPipe; ++u)
            if (u0 + u < cnt) acc += (double)v[u];
        } else {
#pragma unroll
          for (int u = 0; u < kPipe; ++u)
            if (u0 + u < cnt) acc += (double)(v[u] * sc);
        }
      }
    }
  } else if (P.kind == kSparse) {
    const int k0 = uni_i(P.r_ptr[row]), k1 = uni_i(P.r_ptr[row + 1]);
    for (int k = k0; k < k1; k += kPipe) {
      // lanes 0..kPipe-1 fetch the next indices and values; readlane makes
      // them scalar
      const bool ld = lane < kPipe && k + lane < k1;
      const int ind = ld ? P.r_ind[k + lane] : 0;
      const float val = ld ? P.r_val[k + lane] : 0.0f;
      const int cnt = min(kPipe, k1 - k);
      float v[kPipe], w[kPipe];
#pragma unroll
      for (int u = 0; u < kPipe; ++u) {
        const int iu = __builtin_amdgcn_readlane(ind, u);
        w[u] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, val), u));
        v[u] = u < cnt ? (xb + (long long)iu * 64)[lane] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < kPipe; ++u)
        if (u < cnt) acc += (double)w[u] * (double)v[u];
    }
  } else {
    const long long base = (long long)row * gd.n;
    int j = 0;
    if (P.dtype == TV3_F32) {
      const float *A = (const float *)P.A + base;
      for (; j < gd.n; j += kPipe) {
        const int cnt = min(kPipe, gd.n - j);
        const float a = (lane < cnt) ? A[j + lane] : 0.0f;
        float v[kPipe], w[kPipe];
#pragma unroll
        for (int u = 0; u < kPipe; ++u) {
          w[u] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, a), u));
          v[u] = u < cnt ? (xb + (long long)(j + u) * 64)[lane] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kPipe; ++u)
          if (u < cnt) acc += (double)w[u] * (double)v[u];
      }
    } else {
      const __hip_bfloat16 *A = (const __hip_bfloat16 *)P.A + base;
      for (; j < gd.n; j += kPipe) {
        const int cnt = min(kPipe, gd.n - j);
        const float a = (lane < cnt) ? __bfloat162float(A[j + lane]) : 0.0f;
        float v[kPipe], w[kPipe];
#pragma unroll
        for (int u = 0; u < kPipe; ++u) {
          w[u] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, a), u));
          v[u] = u < cnt ? (xb + (long long)(j + u) * 64)[lane] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kPipe; ++u)
          if (u < cnt) acc += (double)w[u] * (double)v[u];
      }
    }
  }
  if (on) y[gd.off_m + (long long)row * 64 + lane] = (float)acc;
}

// W = 64, on-the-fly projector, 16-byte gathers.  One wavefront per ray;
// lane l owns instances 4c .. 4c+3 (c = l % 16) and quarter q = l / 16 takes
// the walk steps s = 4t + q of every 64-step chunk, so one load instruction
// fetches four voxels' 256-byte rows (a 64-lane 4-byte gather fetches one).
// Each lane computes the voxel of one step of the chunk by the closed form;
// quarters read theirs with a lane shuffle.  The four quarter sums are
// combined in a fixed order; FP64 sums of float values, as in k_fwd_w64.
constexpr int kPipeQ = 4;

namespace {
thread_local DenseTune g_dense_tune{2, 2, 8};
}
void dense_tune_set(DenseTune t) { g_dense_tune = t; }
DenseTune dense_tune_cur() { return g_dense_tune; }

// TV3_PROJ_VEC=0 selects the 4-byte-gather kernels (k_fwd_w64 / k_bwd_w64)
// for the projector; both give the same results
bool proj_vec() {
  static const int v = [] {
    const char *e = std::getenv("TV3_PROJ_VEC");
    return (e && e[0] == '0') ? 0 : 1;
  }();
  return v != 0;
}

// Voxel tile of the transpose's column order (TV3_VOX_TILE=t: t x t squares
// of each slice, 0 = raster order).  Default 32, the best of a sweep on the
// 256-transducer ring (transpose 0.194 -> 0.186 ms for two 64-lane groups).
int vox_tile() {
  static const int v = [] {
    const char *e = std::getenv("TV3_VOX_TILE");
    return e ? std::max(0, std::atoi(e)) : 32;
  }();
  return v;
}

// Brick of the transpose's column order for 3-D volumes (TV3_VOX_BRICK=b:
// b x b x b bricks in raster order of bricks; 0 keeps the squares of
// TV3_VOX_TILE in each slice).
int vox_brick() {
  static const int v = [] {
    const char *e = std::getenv("TV3_VOX_BRICK");
    return e ? std::max(0, std::atoi(e)) : 0;
  }();
  return v;
}

// TV3_VOX_NZFIRST=0 keeps the voxels no ray crosses in their tile places of
// the transpose order (default for 3-D volumes: they go last, and the W = 1
// transpose writes them one thread per voxel)
bool vox_nz_first() {
  static const bool v = [] {
    const char *e = std::getenv("TV3_VOX_NZFIRST");
    return !(e && e[0] == '0');
  }();
  return v;
}

// Tile of the forward walk order: rays from te consecutive emitters to tr
// consecutive receivers (TV3_RAY_TILE=te,tr; 0,0 keeps the measurement
// order).  Default 16 x 32, the best of a sweep on the 256-transducer ring
// (forward 0.161 -> 0.135 ms for two 64-lane groups): the rays in flight
// then cross fewer distinct voxels.
void ray_tile(int &te, int &tr) {
  static const std::pair<int, int> v = [] {
    int a = 16, b = 32;
    if (const char *e = std::getenv("TV3_RAY_TILE")) {
      if (std::sscanf(e, "%d,%d", &a, &b) != 2) a = b = 16;
    }
    return std::make_pair(std::max(0, a), std::max(0, b));
  }();
  te = v.first;
  tr = v.second;
}

__device__ __forceinline__ double quarter_sum(double v) {
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

// Row gathers of the 64-lane projector kernels go through a buffer resource
// over the group's vector: a lane's 16 bytes of row r are at byte offset
// 256 r + 16 c.  A step past the end of a walk or list takes the row one
// past the vector (offset >= the resource's size), which the hardware's range
// check returns as zeros: adding +0.0 leaves every sum (which starts at +0.0)
// unchanged, so the rounds need no bounds tests and no masked branches (a
// non-finite scale factor makes the sums non-finite either way).  Byte
// sizes and offsets are 32-bit: batches of geometries with 2^24 or more
// voxels or rays run in W = 1 groups.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rows_rsrc(const float *base, int rows) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)((unsigned)rows * 256u), 0x00020000);
}
__device__ __forceinline__ float4 row_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  return make_float4(__int_as_float(v[0]), __int_as_float(v[1]), __int_as_float(v[2]), __int_as_float(v[3]));
}

// the four instance sums of one lane over rounds t of 64 list or walk steps:
// step 4t + q of the chunk is held, as a row byte offset, by lane 4t + q
template <bool UNIT>
__device__ __forceinline__ void chunk_sums(__amdgpu_buffer_rsrc_t r, int myoff, unsigned coff, int q,
                                           int tmax, float sc, double &a0, double &a1, double &a2,
                                           double &a3) {
  for (int t0 = 0; t0 < tmax; t0 += kPipeQ) {
    float4 v[kPipeQ];
#pragma unroll
    for (int u = 0; u < kPipeQ; ++u) {
      const int off = __builtin_amdgcn_ds_bpermute(16 * (t0 + u) + 4 * q, myoff);
      v[u] = row_load(r, (unsigned)off + coff);
    }
#pragma unroll
    for (int u = 0; u < kPipeQ; ++u) {
      if (UNIT) {
        a0 += (double)v[u].x; a1 += (double)v[u].y;
        a2 += (double)v[u].z; a3 += (double)v[u].w;
      } else {
        a0 += (double)(v[u].x * sc); a1 += (double)(v[u].y * sc);
        a2 += (double)(v[u].z * sc); a3 += (double)(v[u].w * sc);
      }
    }
  }
}

template <bool UNIT>
__device__ __forceinline__ void fwd_ray64(const RayDesc &rd, int L, __amdgpu_buffer_rsrc_t r, int n,
                                          unsigned coff, int q, int lane, float sc, double &a0, double &a1,
                                          double &a2, double &a3) {
  for (int j0 = 0; j0 <= L; j0 += 64) {
    const int jl = j0 + lane;
    const int myoff = (int)((unsigned)(jl <= L ? ray_lin(rd, jl) : n) * 256u);
    const int cnt = min(64, L + 1 - j0);
    chunk_sums<UNIT>(r, myoff, coff, q, (cnt + 3) >> 2, sc, a0, a1, a2, a3);
  }
}

__global__ __launch_bounds__(kBlock) void k_fwd_w64q(const GroupDesc *G,
                                                     const float *x, float *y,
                                                     const InstScal *S,
                                                     uint32_t mask) {
  const GroupDesc &gd = G[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, c = lane & 15;
  const int slot = uni_i(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
  if (slot >= gd.m) return;
  const int row = gd.plan.ray_order ? uni_i(ldg(gd.plan.ray_order + slot)) : slot;
  const InstScal *Sg = S + blockIdx.y * 64 + 4 * c;
  bool on[4];
  bool any = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    on[k] = !mask || (Sg[k].flags & mask) == mask;
    any = any || on[k];
  }
  if (!__any(any)) return;
  const PlanDev &P = gd.plan;
  const RayDesc rd = ldg(P.rays + row);
  const int L = uni_i(rd.L);
  const float sc = P.scale;
  const int n = gd.n;
  const __amdgpu_buffer_rsrc_t xr = rows_rsrc(x + gd.off_n, n);
  const unsigned coff = 16u * c;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (sc == 1.0f)
    fwd_ray64<true>(rd, L, xr, n, coff, q, lane, sc, a0, a1, a2, a3);
  else
    fwd_ray64<false>(rd, L, xr, n, coff, q, lane, sc, a0, a1, a2, a3);
  a0 = quarter_sum(a0); a1 = quarter_sum(a1);
  a2 = quarter_sum(a2); a3 = quarter_sum(a3);
  if (q == 0) {
    float *yo = y + gd.off_m + (long long)row * 64 + 4 * c;
    if (on[0] && on[1] && on[2] && on[3]) {
      *(float4 *)yo = make_float4((float)a0, (float)a1, (float)a2, (float)a3);
    } else {
      if (on[0]) yo[0] = (float)a0;
      if (on[1]) yo[1] = (float)a1;
      if (on[2]) yo[2] = (float)a2;
      if (on[3]) yo[3] = (float)a3;
    }
  }
}

// A^T of the on-the-fly projector with 16-byte gathers.  A wavefront walks
// kBwdCols consecutive voxels of the plan's transpose order (t_col; the ray
// lists t_ray[t_ptr[s] .. t_ptr[s+1]) ascending); quarters take list entries
// 4t + q of each 64-entry chunk, lanes own four instances (as k_fwd_w64q).
// The list bounds of the voxel after next and the first index chunk of the
// next voxel are loaded while the current voxel's gathers are in flight
// (two voxels per wavefront measured best: 0.186 -> 0.173 ms for two 64-lane
// groups of the 256 ring; four 0.186, eight 0.237).
constexpr int kBwdCols = 2;

template <bool UNIT>
__device__ __forceinline__ void bwd_list64(const int *t_ray, int kA0, int kA1, int idA,
                                           __amdgpu_buffer_rsrc_t r, int m, unsigned coff, int q,
                                           int lane, float sc, double &a0, double &a1, double &a2,
                                           double &a3) {
  for (int kb = kA0; kb < kA1; kb += 64) {
    const int cnt = min(64, kA1 - kb);
    const int id = kb == kA0 ? idA : (lane < cnt ? ldg(t_ray + kb + lane) : m);
    chunk_sums<UNIT>(r, (int)((unsigned)id * 256u), coff, q, (cnt + 3) >> 2, sc, a0, a1, a2, a3);
  }
}

__global__ __launch_bounds__(kBlock) void k_bwd_w64q(const GroupDesc *G,
                                                     const float *x, float *y,
                                                     const InstScal *S,
                                                     uint32_t mask,
                                                     const float *sub) {
  const GroupDesc &gd = G[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, c = lane & 15;
  const int s0 = uni_i((blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * kBwdCols);
  if (s0 >= gd.n) return;
  const int ns = min(kBwdCols, gd.n - s0);
  const InstScal *Sg = S + blockIdx.y * 64 + 4 * c;
  bool on[4];
  bool any = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    on[k] = !mask || (Sg[k].flags & mask) == mask;
    any = any || on[k];
  }
  if (!__any(any)) return;
  const PlanDev &P = gd.plan;
  const float sc = P.scale;
  const bool unit = sc == 1.0f;
  const int m = gd.m;
  const __amdgpu_buffer_rsrc_t xr = rows_rsrc(x + gd.off_m, m);
  const unsigned coff = 16u * c;
  // voxel A: current, B: next (bounds known), C: after next
  int kA0 = uni_i(ldg(P.t_ptr + (s0))), kA1 = uni_i(ldg(P.t_ptr + (s0 + 1)));
  int colA = uni_i(ldg(P.t_col + (s0)));
  int idA = lane < min(64, kA1 - kA0) ? ldg(P.t_ray + (kA0 + lane)) : m;
  int kB0 = 0, kB1 = 0, colB = 0;
  if (ns > 1) {
    kB0 = uni_i(ldg(P.t_ptr + (s0 + 1)));
    kB1 = uni_i(ldg(P.t_ptr + (s0 + 2)));
    colB = uni_i(ldg(P.t_col + (s0 + 1)));
  }
  for (int i = 0; i < ns; ++i) {
    const int idB = (i + 1 < ns && lane < min(64, kB1 - kB0)) ? ldg(P.t_ray + (kB0 + lane)) : m;
    int kC0 = 0, kC1 = 0, colC = 0;
    if (i + 2 < ns) {
      kC0 = uni_i(ldg(P.t_ptr + (s0 + i + 2)));
      kC1 = uni_i(ldg(P.t_ptr + (s0 + i + 3)));
      colC = uni_i(ldg(P.t_col + (s0 + i + 2)));
    }
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (unit)
      bwd_list64<true>(P.t_ray, kA0, kA1, idA, xr, m, coff, q, lane, sc, a0, a1, a2, a3);
    else
      bwd_list64<false>(P.t_ray, kA0, kA1, idA, xr, m, coff, q, lane, sc, a0, a1, a2, a3);
    a0 = quarter_sum(a0); a1 = quarter_sum(a1);
    a2 = quarter_sum(a2); a3 = quarter_sum(a3);
    if (q == 0) {
      const long long o = gd.off_n + (long long)colA * 64 + 4 * c;
      float r[4] = {(float)a0, (float)a1, (float)a2, (float)a3};
      if (sub) {
#pragma unroll
        for (int k = 0; k < 4; ++k) r[k] = r[k] - sub[o + k];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (on[k]) y[o + k] = r[k];
    }
    kA0 = kB0; kA1 = kB1; colA = colB; idA = idB;
    kB0 = kC0; kB1 = kC1; colB = colC;
  }
}

// ---------------------------------------------------------------------------
// transposed product: y[col] = sum_k A[k, col] x[k] (- sub[col])
// ---------------------------------------------------------------------------

// A geometry plan of one layer sends the voxels no ray crosses (slots
// t_nz .. n of its transpose order) to blocks of their own, one thread per
// voxel, after the ceil(t_nz / (kBlock / L)) blocks of L lanes per voxel; the
// grid has one block more than ceil(n / (kBlock / L)), which covers both.
template <int L>
__global__ __launch_bounds__(kBlock) void k_bwd_w1(const GroupDesc *G,
                                                   const float *x, float *y,
                                                   const InstScal *S,
                                                   uint32_t mask,
                                                   const float *sub) {
  const GroupDesc &gd = G[blockIdx.y];
  const int sl = threadIdx.x & (L - 1);
  const int col = blockIdx.x * (kBlock / L) + (threadIdx.x / L);
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const PlanDev &P = gd.plan;
  // column of A and layer of this output entry (one layer: the column
  // itself; a geometry column slot walks the plan's tiled voxel order)
  const bool geo = P.kind == kGeometry;
  const bool tiled = kW1TiledT && geo && P.t_col;
  const bool split = tiled && gd.n == P.n;
  if (split) {
    const int nbA = (P.t_nz + kBlock / L - 1) / (kBlock / L);
    if ((int)blockIdx.x >= nbA) {
      const long long k = P.t_nz + (long long)(blockIdx.x - nbA) * kBlock + threadIdx.x;
      if (k < P.n) {
        const long long o = gd.off_n + ldg(P.t_col + k);
        float v = 0.0f;
        if (sub) v = v - sub[o];
        y[o] = v;
      }
      return;
    }
  }
  const int colEnd = split ? P.t_nz : gd.n;
  int ac = col % P.n;
  const float *xg = x + gd.off_m + (long long)(col / P.n) * P.m;
  double acc = 0.0;
  if (col < colEnd) {
    constexpr int U = kW1Unroll;
    const int *ptr = tiled ? P.t_ptr : geo ? P.g_cptr : P.c_ptr;
    const int *ind = tiled ? P.t_ray : geo ? P.g_cray : P.c_ind;
    const int k0 = ldg(ptr + ac), k1 = ldg(ptr + ac + 1);
    if (tiled) ac = ldg(P.t_col + col);
    const float sc = P.scale;
    for (int k = k0 + sl; k < k1; k += L * U) {
      int id[U];
      float w[U], v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kk = k + u * L;
        id[u] = kk < k1 ? ldg(ind + kk) : -1;
        w[u] = (!geo && kk < k1) ? ldg(P.c_val + kk) : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = id[u] >= 0 ? xg[id[u]] : 0.0f;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (id[u] >= 0) acc += geo ? (double)(v[u] * sc) : (double)w[u] * (double)v[u];
    }
  }
  acc = lane_sum<L>(acc);
  if (sl == 0 && col < colEnd) {
    const long long o = gd.off_n + (long long)(col / P.n) * P.n + ac;
    float v = (float)acc;
    if (sub) v = v - sub[o];
    y[o] = v;
  }
}

// Staged transpose of a geometry plan (PlanDev::s_*).  A block takes one tile
// of kStageV voxel slots, bricks of the volume.  It gathers the y entries of
// the tile's rays into LDS, one load per ray instead of one per list entry
// (a ray through a brick crosses several of its voxels: on the 128^3
// hemisphere 8.5 M loads replace 32.4 M), kStageLd loads in flight per
// thread.  L lanes per voxel then sum the voxel's list from LDS as k_bwd_w1
// sums it (lane sl takes entries sl, sl + L, ... in order, the same
// butterfly adds the lanes).  The list bounds of all kStageP passes are
// loaded up front, and the first NI entries per lane of the next pass's
// voxel load while the current pass sums, so a pass waits for memory about
// once.  A tile crossed by more than NR rays gathers from y directly (its
// voxels' lists in g_cray).  Lists longer than kStageH entries (the voxels
// at the transducers: about a thousand rays each on the hemisphere) would
// hold their wavefront for dozens of dependent loads; they go to
// kHeavyBlocks blocks instead, a wavefront per voxel with 64 lanes on its
// list.  The blocks after those write the voxels no ray crosses, kBlock per
// block, so ceil(max_n / kBlock) + 1 + kHeavyBlocks blocks per group cover
// everything.  Sums are FP64 and rounded once, as in every other transpose.
// (Geometry plans take no layers: gd.n == P.n.)
constexpr int kStageR = 8192;                          // most rays staged per tile (32 KiB of LDS)
constexpr int kStageLd = 8;                            // staging loads in flight per thread
constexpr int kStageV = kBlock;                        // voxel slots per tile
constexpr int kStageP = kStageV / (kBlock / kSubT);    // passes per tile
constexpr int kStageI = 16;                            // most entries per lane loaded ahead
constexpr int kStageH = kStageI * kSubT;               // longest list of a tile slot
constexpr int kHeavyBlocks = 512;                      // blocks that take the longer lists

// list bounds of the kStageP passes and the first entries of pass 0 (issued
// before the staging loads, so they are in flight together)
template <int NI>
struct StagePre {
  int vv[kStageP], kb[kStageP], ke[kStageP];
  int cur[NI];
};

template <int L, int NI, bool ST, class I>
__device__ __forceinline__ void stage_pre(StagePre<NI> &q, const int *scol, const int *ptr, const I *ind,
                                          int s_beg, int s_end) {
  const int sl = threadIdx.x & (L - 1);
  const int s_me = s_beg + (int)threadIdx.x / L;
#pragma unroll
  for (int p = 0; p < kStageP; ++p) {
    const int s = s_me + p * (kBlock / L);
    q.vv[p] = s < s_end ? ldg(scol + s) : 0;
  }
#pragma unroll
  for (int p = 0; p < kStageP; ++p) {
    const int s = s_me + p * (kBlock / L);
    const int i = ST ? s : q.vv[p];
    q.kb[p] = s < s_end ? ldg(ptr + i) : 0;
    q.ke[p] = s < s_end ? ldg(ptr + i + 1) : 0;
  }
#pragma unroll
  for (int j = 0; j < NI; ++j) {
    const int k = q.kb[0] + sl + j * L;
    q.cur[j] = k < q.ke[0] ? (int)ldg(ind + k) : -1;
  }
}

template <int L, int NI, class I, class F>
__device__ __forceinline__ void stage_sums(StagePre<NI> &q, const I *ind, F val, int s_beg, int s_end,
                                           float sc, float *yg, const float *subg) {
  static_assert(L == kSubT, "tile passes assume kSubT lanes per voxel");
  constexpr int U = kW1Unroll;
  const int sl = threadIdx.x & (L - 1);
  const int s_me = s_beg + (int)threadIdx.x / L;
  int nxt[NI];
#pragma unroll
  for (int p = 0; p < kStageP; ++p) {
    if (s_beg + p * (kBlock / L) >= s_end) break;  // uniform: the tile's last pass
    if (p + 1 < kStageP) {
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        const int k = q.kb[p + 1] + sl + j * L;
        nxt[j] = k < q.ke[p + 1] ? (int)ldg(ind + k) : -1;
      }
    }
    double acc = 0.0;
    {
      float w[NI];
#pragma unroll
      for (int j = 0; j < NI; ++j) w[j] = q.cur[j] >= 0 ? val(q.cur[j]) : 0.0f;
#pragma unroll
      for (int j = 0; j < NI; ++j)
        if (q.cur[j] >= 0) acc += (double)(w[j] * sc);
    }
    for (int k = q.kb[p] + sl + NI * L; k < q.ke[p]; k += L * U) {
      int id[U];
      float w[U];
#pragma unroll
      for (int u = 0; u < U; ++u) id[u] = k + u * L < q.ke[p] ? (int)ldg(ind + k + u * L) : -1;
#pragma unroll
      for (int u = 0; u < U; ++u) w[u] = id[u] >= 0 ? val(id[u]) : 0.0f;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (id[u] >= 0) acc += (double)(w[u] * sc);
    }
    acc = lane_sum<L>(acc);
    if (sl == 0 && s_me + p * (kBlock / L) < s_end) {
      const int v = q.vv[p];
      float r = (float)acc;
      if (subg) r = r - ldg(subg + v);
      yg[v] = r;
    }
    if (p + 1 < kStageP) {
#pragma unroll
      for (int j = 0; j < NI; ++j) q.cur[j] = nxt[j];
    }
  }
}

template <int L, int NI, int NR>
__global__ __launch_bounds__(kBlock) void k_bwd_w1s(const GroupDesc *G,
                                                    const float *x, float *y,
                                                    const InstScal *S,
                                                    uint32_t mask,
                                                    const float *sub) {
  const GroupDesc &gd = G[blockIdx.y];
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const PlanDev &P = gd.plan;
  const int n = P.n;
  const int s_nz = P.s_nz;
  const int T = (s_nz + kStageV - 1) / kStageV;
  const int t = blockIdx.x;
  float *yg = y + gd.off_n;
  const float *subg = sub ? sub + gd.off_n : nullptr;
  const float sc = P.scale;
  const float *xg = x + gd.off_m;
  if (t >= T + kHeavyBlocks) {
    const long long k = s_nz + P.s_nh + (long long)(t - T - kHeavyBlocks) * kBlock + threadIdx.x;
    if (k >= n) return;
    const int v = ldg(P.s_col + k);
    float r = 0.0f;
    if (subg) r = r - ldg(subg + v);
    yg[v] = r;
    return;
  }
  if (t >= T) {
    // the voxels with lists longer than kStageH, longest first: a wavefront
    // per voxel, lane l takes entries l, l + 64, ..., the wave's lanes are
    // added by a butterfly
    const int lane = threadIdx.x & 63;
    for (int h = (t - T) * (kBlock / 64) + (int)threadIdx.x / 64; h < P.s_nh;
         h += kHeavyBlocks * (kBlock / 64)) {
      const int v = ldg(P.s_col + s_nz + h);
      const int k0 = ldg(P.g_cptr + v), k1 = ldg(P.g_cptr + v + 1);
      double acc = 0.0;
      for (int k = k0 + lane; k < k1; k += 64 * kStageI) {
        int id[kStageI];
        float w[kStageI];
#pragma unroll
        for (int u = 0; u < kStageI; ++u) id[u] = k + 64 * u < k1 ? ldg(P.g_cray + k + 64 * u) : -1;
#pragma unroll
        for (int u = 0; u < kStageI; ++u) w[u] = id[u] >= 0 ? ldg(xg + id[u]) : 0.0f;
#pragma unroll
        for (int u = 0; u < kStageI; ++u)
          if (id[u] >= 0) acc += (double)(w[u] * sc);
      }
      acc = lane_sum<64>(acc);
      if (lane == 0) {
        float r = (float)acc;
        if (subg) r = r - ldg(subg + v);
        yg[v] = r;
      }
    }
    return;
  }
  __shared__ float ys[NR];
  const int r0 = ldg(P.s_rptr + t), r1 = ldg(P.s_rptr + t + 1);
  const int s_beg = t * kStageV, s_end = min(s_nz, s_beg + kStageV);
  StagePre<NI> q;
  if (r1 > r0) {
    stage_pre<L, NI, true>(q, P.s_col, P.s_ptr, P.s_loc, s_beg, s_end);
    for (int i0 = r0 + (int)threadIdx.x; i0 < r1; i0 += kBlock * kStageLd) {
      int id[kStageLd];
      float w[kStageLd];
#pragma unroll
      for (int u = 0; u < kStageLd; ++u) id[u] = i0 + u * kBlock < r1 ? ldg(P.s_ray + i0 + u * kBlock) : -1;
#pragma unroll
      for (int u = 0; u < kStageLd; ++u) w[u] = id[u] >= 0 ? ldg(xg + id[u]) : 0.0f;
#pragma unroll
      for (int u = 0; u < kStageLd; ++u)
        if (id[u] >= 0) ys[i0 + u * kBlock - r0] = w[u];
    }
    __syncthreads();
    stage_sums<L>(q, P.s_loc, [&](int i) { return ys[i]; }, s_beg, s_end, sc, yg, subg);
  } else {
    stage_pre<L, NI, false>(q, P.s_col, P.g_cptr, P.g_cray, s_beg, s_end);
    stage_sums<L>(q, P.g_cray, [&](int i) { return ldg(xg + i); }, s_beg, s_end, sc, yg, subg);
  }
}

// staged transpose shape: list entries per lane loaded ahead
// (TV3_STAGE_DEPTH = 8, 12, 16) and the most rays a tile stages
// (TV3_STAGE_RAYS = 6144 or 8192, read when a plan is built); neither
// changes a result
int stage_depth() {
  static const int v = [] {
    const char *e = std::getenv("TV3_STAGE_DEPTH");
    const int d = e ? std::atoi(e) : 0;
    return d > 0 ? d : kStageI;
  }();
  return v;
}
int stage_rays() {
  static const int v = [] {
    const char *e = std::getenv("TV3_STAGE_RAYS");
    const int r = e ? std::atoi(e) : 0;
    return r > 0 && r <= 6144 ? 6144 : kStageR;
  }();
  return v;
}

// TV3_STAGE=0 keeps the unstaged single-instance transpose (k_bwd_w1) for
// geometry plans; both give the same results
bool stage_enabled() {
  static const bool v = [] {
    const char *e = std::getenv("TV3_STAGE");
    return !(e && e[0] == '0');
  }();
  return v;
}

__global__ __launch_bounds__(kBlock) void k_bwd_w64(const GroupDesc *G,
                                                    const float *x, float *y,
                                                    const InstScal *S,
                                                    uint32_t mask,
                                                    const float *sub,
                                                    int submat_mb) {
  const GroupDesc &gd = G[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int col = uni_i(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
  if (col >= gd.n) return;
  const bool on = !mask || (S[blockIdx.y * 64 + lane].flags & mask) == mask;
  if (!__any(on)) return;
  const PlanDev &P = gd.plan;
  const float *xb = x + gd.off_m;
  double acc = 0.0;
  if (P.kind == kGeometry || P.kind == kSparse) {
    const bool geo = P.kind == kGeometry;
    const int *ptr = geo ? P.g_cptr : P.c_ptr;
    const int *ind = geo ? P.g_cray : P.c_ind;
    const float sc = P.scale;
    const bool unit = geo && sc == 1.0f;
    const int k0 = uni_i(ptr[col]), k1 = uni_i(ptr[col + 1]);
    for (int kb = k0; kb < k1; kb += 64) {
      const int cnt64 = min(64, k1 - kb);
      const bool ld = lane < cnt64;
      const int id = ld ? ind[kb + lane] : 0;
      const float val = (ld && !geo) ? P.c_val[kb + lane] : sc;
      for (int u0 = 0; u0 < cnt64; u0 += kPipe) {
        float v[kPipe], w[kPipe];
#pragma unroll
        for (int u = 0; u < kPipe; ++u) {
          const int iu = __builtin_amdgcn_readlane(id, u0 + u);
          w[u] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, val), u0 + u));
          v[u] = (u0 + u < cnt64) ? (xb + (long long)iu * 64)[lane] : 0.0f;
        }
        if (unit) {
#pragma unroll
          for (int u = 0; u < kPipe; ++u)
            if (u0 + u < cnt64) acc += (double)v[u];
        } else if (geo) {
#pragma unroll
          for (int u = 0; u < kPipe; ++u)
            if (u0 + u < cnt64) acc += (double)(v[u] * sc);
        } else {
#pragma unroll
          for (int u = 0; u < kPipe; ++u)
            if (u0 + u < cnt64) acc += (double)w[u] * (double)v[u];
        }
      }
    }
  } else {
    // dense: rows in order; with submat_mb the FP64 sum restarts at every
    // block boundary and the rounded block sums are added in float
    const int rb = submat_mb > 0 ? dense_block_rows(submat_mb, gd.n) : 0x7fffffff;
    int next_b = rb;
    float gf = 0.0f;
    for (int i = 0; i < gd.m; i += kPipe) {
      const int cnt = min(kPipe, gd.m - i);
      float a = 0.0f;
      if (lane < cnt) {
        const long long ia = (long long)(i + lane) * gd.n + col;
        a = P.dtype == TV3_F32 ? ((const float *)P.A)[ia]
                               : __bfloat162float(((const __hip_bfloat16 *)P.A)[ia]);
      }
      float v[kPipe], w[kPipe];
#pragma unroll
      for (int u = 0; u < kPipe; ++u) {
        w[u] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, a), u));
        v[u] = u < cnt ? (xb + (long long)(i + u) * 64)[lane] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < kPipe; ++u) {
        if (u < cnt) {
          if (i + u == next_b) {
            gf = gf + (float)acc;
            acc = 0.0;
            next_b += rb;
          }
          acc += (double)w[u] * (double)v[u];
        }
      }
    }
    if (submat_mb > 0) {
      gf = gf + (float)acc;
      acc = (double)gf;  // exact: gf is a float
    }
  }
  if (on) {
    const long long o = gd.off_n + (long long)col * 64 + lane;
    float v = (float)acc;
    if (sub) v = v - sub[o];
    y[o] = v;
  }
}

// ---------------------------------------------------------------------------
// W = 16, XCD-affine (shared-geometry batches)
// ---------------------------------------------------------------------------
// A group of 16 instances keeps its image (n x 16 floats) and measurements
// (m x 16 floats) in one XCD's L2: with at least eight groups, the flat block
// index b runs group (b % 8) + 8 * slot, so every block of a group lands on
// the same XCD (workgroups are dealt to the eight XCDs round robin).  A
// wavefront handles one row; its four 16-lane quarters take four consecutive
// entries per load (4 x 64 bytes), kQuad loads in flight; the quarter sums
// are combined by a fixed butterfly.
constexpr int kXcd = 8;
constexpr int kQuad = 4;

__device__ __forceinline__ bool grp_decode(int nG, int RB, int affine, int &g,
                                           int &rb) {
  if (!affine) {
    g = blockIdx.y;
    rb = blockIdx.x;
    return true;
  }
  const int b = blockIdx.x;
  const int t = b / kXcd;
  g = (b % kXcd) + kXcd * (t / RB);
  rb = t % RB;
  return g < nG;
}

__global__ __launch_bounds__(kBlock) void k_fwd_w16(const GroupDesc *G, int nG,
                                                    int RB, int affine,
                                                    const float *x, float *y,
                                                    const InstScal *S,
                                                    uint32_t mask) {
  int g, rb;
  if (!grp_decode(nG, RB, affine, g, rb)) return;
  const GroupDesc &gd = G[g];
  const int lane = threadIdx.x & 63, q4 = lane >> 4, bl = lane & 15;
  const int row = uni_i(rb * (kBlock / 64) + (threadIdx.x >> 6));
  if (row >= gd.m) return;
  const bool on = !mask || (S[g * 16 + bl].flags & mask) == mask;
  if (!__any(on)) return;
  const PlanDev &P = gd.plan;
  const float *xb = x + gd.off_n + bl;
  double acc = 0.0;
  if (P.kind == kGeometry) {
    // 64 walk steps per round, one index per lane (closed form); quarter q4
    // of the wavefront takes steps j0 + 4u + q4 through a lane shuffle
    const RayDesc rd = P.rays[row];
    const int L = uni_i(rd.L);
    const float sc = P.scale;
    for (int j0 = 0; j0 <= L; j0 += 64) {
      const int jl = j0 + lane;
      const int myidx = jl <= L ? ray_lin(rd, jl) : 0;
      for (int u0 = 0; u0 < 16; u0 += kQuad) {
        if (j0 + 4 * u0 > L) break;
        float v[kQuad];
#pragma unroll
        for (int q = 0; q < kQuad; ++q) {
          const int t = 4 * (u0 + q) + q4;
          const int iu = __shfl(myidx, t, 64);
          v[q] = (j0 + t <= L) ? xb[(long long)iu * 16] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < kQuad; ++q)
          if (j0 + 4 * (u0 + q) + q4 <= L) acc += (double)(v[q] * sc);
      }
    }
  } else {
    const int k0 = uni_i(P.r_ptr[row]), k1 = uni_i(P.r_ptr[row + 1]);
    for (int k = k0; k < k1; k += 4 * kQuad) {
      float v[kQuad], w[kQuad];
#pragma unroll
      for (int q = 0; q < kQuad; ++q) {
        const int kk = k + 4 * q + q4;
        const bool ok = kk < k1;
        w[q] = ok ? P.r_val[kk] : 0.0f;
        v[q] = ok ? xb[(long long)P.r_ind[kk] * 16] : 0.0f;
      }
#pragma unroll
      for (int q = 0; q < kQuad; ++q)
        if (k + 4 * q + q4 < k1) acc += (double)w[q] * (double)v[q];
    }
  }
  acc += __shfl_xor(acc, 16, 64);
  acc += __shfl_xor(acc, 32, 64);
  if (on && q4 == 0) y[gd.off_m + (long long)row * 16 + bl] = (float)acc;
}

__global__ __launch_bounds__(kBlock) void k_bwd_w16(const GroupDesc *G, int nG,
                                                    int RB, int affine,
                                                    const float *x, float *y,
                                                    const InstScal *S,
                                                    uint32_t mask,
                                                    const float *sub) {
  int g, rb;
  if (!grp_decode(nG, RB, affine, g, rb)) return;
  const GroupDesc &gd = G[g];
  const int lane = threadIdx.x & 63, q4 = lane >> 4, bl = lane & 15;
  const int col = uni_i(rb * (kBlock / 64) + (threadIdx.x >> 6));
  if (col >= gd.n) return;
  const bool on = !mask || (S[g * 16 + bl].flags & mask) == mask;
  if (!__any(on)) return;
  const PlanDev &P = gd.plan;
  const bool geo = P.kind == kGeometry;
  const int *ptr = geo ? P.g_cptr : P.c_ptr;
  const int *ind = geo ? P.g_cray : P.c_ind;
  const float sc = P.scale;
  const float *xb = x + gd.off_m + bl;
  double acc = 0.0;
  const int k0 = uni_i(ptr[col]), k1 = uni_i(ptr[col + 1]);
  for (int kb = k0; kb < k1; kb += 64) {
    const int cnt64 = min(64, k1 - kb);
    const bool ld = lane < cnt64;
    const int id = ld ? ind[kb + lane] : 0;
    const float val = (ld && !geo) ? P.c_val[kb + lane] : 0.0f;
    for (int u0 = 0; u0 < 16; u0 += kQuad) {
      if (4 * u0 >= cnt64) break;
      float v[kQuad], w[kQuad];
#pragma unroll
      for (int q = 0; q < kQuad; ++q) {
        const int t = 4 * (u0 + q) + q4;
        const int iu = __shfl(id, t, 64);
        w[q] = __shfl(val, t, 64);
        v[q] = (t < cnt64) ? xb[(long long)iu * 16] : 0.0f;
      }
#pragma unroll
      for (int q = 0; q < kQuad; ++q)
        if (4 * (u0 + q) + q4 < cnt64)
          acc += geo ? (double)(v[q] * sc) : (double)w[q] * (double)v[q];
    }
  }
  acc += __shfl_xor(acc, 16, 64);
  acc += __shfl_xor(acc, 32, 64);
  if (on && q4 == 0) {
    const long long o = gd.off_n + (long long)col * 16 + bl;
    float r = (float)acc;
    if (sub) r = r - sub[o];
    y[o] = r;
  }
}

// ---------------------------------------------------------------------------
// dense A with W = 1
// ---------------------------------------------------------------------------
// A entry -> float (bf16 widens exactly)
__device__ __forceinline__ float a_val(const float *A, long long i) { return A[i]; }
__device__ __forceinline__ float a_val(const __hip_bfloat16 *A, long long i) {
  return __bfloat162float(A[i]);
}
// eight consecutive entries from a 16-byte (bf16) / 32-byte (f32) aligned
// address as one or two vector loads
__device__ __forceinline__ void load8(const float *p, float (&v)[8]) {
  const float4 a = *(const float4 *)p, b = *(const float4 *)(p + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void load8(const __hip_bfloat16 *p, float (&v)[8]) {
  const uint4 q = *(const uint4 *)p;
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = __uint_as_float(w[k] << 16);
    v[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
  }
}

// four consecutive entries (8 bytes of bf16, 16 of f32)
__device__ __forceinline__ void load4(const float *p, float (&v)[4]) {
  const float4 a = *(const float4 *)p;
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void load4(const __hip_bfloat16 *p, float (&v)[4]) {
  const uint2 q = *(const uint2 *)p;
  v[0] = __uint_as_float(q.x << 16);
  v[1] = __uint_as_float(q.x & 0xffff0000u);
  v[2] = __uint_as_float(q.y << 16);
  v[3] = __uint_as_float(q.y & 0xffff0000u);
}
// two consecutive entries (4 bytes of bf16, 8 of f32)
__device__ __forceinline__ void load2(const float *p, float (&v)[2]) {
  const float2 a = *(const float2 *)p;
  v[0] = a.x; v[1] = a.y;
}
__device__ __forceinline__ void load2(const __hip_bfloat16 *p, float (&v)[2]) {
  const uint32_t q = *(const uint32_t *)p;
  v[0] = __uint_as_float(q << 16);
  v[1] = __uint_as_float(q & 0xffff0000u);
}
template <int C, class T>
__device__ __forceinline__ void loadc(const T *p, float (&v)[C]) {
  if constexpr (C == 8) load8(p, v); else if constexpr (C == 4) load4(p, v); else load2(p, v);
}

// acc + x y for float-valued x and y (float or bf16 entries, float vector
// values widened to FP64): their product has at most 48 significant bits and
// is exact in FP64, so one fused rounding gives the bits of the product
// rounded, then added (the oracle's order), in one instruction
__device__ __forceinline__ double prod_add(double acc, double x, double y) { return __builtin_fma(x, y, acc); }

// y = A x: one wavefront per R consecutive rows; with n a multiple of 8 every
// lane loads 8 consecutive entries of each row (16 bytes of bf16, 32 of f32)
// and the matching 8 x values once for all R rows per step; FP64 lane sums,
// fixed butterfly.  Each row's summation order is the same for every R, so
// R is a pure speed choice (picked by the autotuner).  With several layers,
// output row l * plan.m + i is row i of A times layer l of x; R rows that
// cross a layer boundary are taken one at a time.
template <class T, int R>
__global__ __launch_bounds__(kBlock) void k_dense_fwd_w1(const GroupDesc *G,
                                                         const float *x,
                                                         float *y,
                                                         const InstScal *S,
                                                         uint32_t mask) {
  const GroupDesc &gd = G[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * R;
  if (row0 >= gd.m) return;
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const int n = gd.plan.n, pm = gd.plan.m;
  const int nr = min(R, gd.m - row0);
  if (R > 1 && row0 / pm != (row0 + nr - 1) / pm) {
    // rows one at a time, each with the lane partition of the main path
    // below (8 consecutive entries per lane with vector rows, else one per
    // lane), so that every row is summed in the same order for every R
    for (int r = 0; r < nr; ++r) {
      const int row = row0 + r;
      const T *Ar = (const T *)gd.plan.A + (long long)(row % pm) * n;
      const float *xr = x + gd.off_n + (long long)(row / pm) * n;
      double a = 0.0;
      if (gd.plan.a_vec) {
        for (int j = lane * 8; j < n; j += 512) {
          float v[8], av[8];
          load8(xr + j, v);
          load8(Ar + j, av);
#pragma unroll
          for (int u = 0; u < 8; ++u) a = prod_add(a, (double)av[u], (double)v[u]);
        }
      } else {
        for (int j = lane; j < n; j += 64) a += (double)a_val(Ar, j) * (double)xr[j];
      }
      a = wave_sum(a);
      if (lane == 0) y[gd.off_m + row] = (float)a;
    }
    return;
  }
  const T *A = (const T *)gd.plan.A + (long long)(row0 % pm) * n;
  const float *xg = x + gd.off_n + (long long)(row0 / pm) * n;
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  // rows start 16-byte aligned when n % 8 == 0; x starts 32-byte aligned
  // (group offsets are multiples of 64 floats)
  if (gd.plan.a_vec) {
    for (int j = lane * 8; j < n; j += 512) {
      float v[8];
      load8(xg + j, v);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r < nr) {
          float a[8];
          load8(A + (long long)r * n + j, a);
#pragma unroll
          for (int u = 0; u < 8; ++u) acc[r] = prod_add(acc[r], (double)a[u], (double)v[u]);
        }
      }
    }
  } else {
    for (int j = lane; j < n; j += 64) {
      const double xv = (double)xg[j];
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r < nr) acc[r] += (double)a_val(A + (long long)r * n, j) * xv;
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double t = wave_sum(acc[r]);
    if (lane == 0 && r < nr) y[gd.off_m + row0 + r] = (float)t;
  }
}

// A^T x, split-K: chunk c of a group covers rows [c*R, min(m, (c+1)*R)); a
// thread owns 8 consecutive columns (one 16/32-byte load per row) and writes
// FP64 partial sums; k_dense_bwd_reduce adds the chunks in order
// (kDenseChunks chunks at most, tv3_internal.h).

template <class T, int U>
__global__ __launch_bounds__(kBlock) void k_dense_bwd_part(
    const GroupDesc *G, const float *x, double *scratch, const InstScal *S,
    uint32_t mask, long long sstride) {
  const GroupDesc &gd = G[blockIdx.z];
  // c0 indexes the layered output (gd.n = layers * n entries); with n % 8 ==
  // 0 (a_vec) a thread's 8 columns lie in one layer
  const int n = gd.plan.n, pm = gd.plan.m;
  const int c0 = (blockIdx.x * kBlock + threadIdx.x) * 8;
  if (c0 >= gd.n) return;
  if (mask && (S[blockIdx.z].flags & mask) != mask) return;
  const int nch = min(kDenseChunks, (pm + kBlock - 1) / kBlock);
  const int c = blockIdx.y;
  if (c >= nch) return;
  const int R = (pm + nch - 1) / nch;
  const int i0 = c * R, i1 = min(pm, i0 + R);
  const T *A = (const T *)gd.plan.A;
  const int w = min(8, gd.n - c0);
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (w == 8 && gd.plan.a_vec) {
    const float *xg = x + gd.off_m + (long long)(c0 / n) * pm;
    const int j0 = c0 % n;
    // U rows loaded before they are consumed, in row order: the per-column
    // summation order does not depend on U
    int i = i0;
    for (; i + U <= i1; i += U) {
      float a[U][8];
      double xi[U];
#pragma unroll
      for (int r = 0; r < U; ++r) {
        load8(A + (long long)(i + r) * n + j0, a[r]);
        xi[r] = (double)xg[i + r];
      }
#pragma unroll
      for (int r = 0; r < U; ++r)
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] = prod_add(acc[u], (double)a[r][u], xi[r]);
    }
    for (; i < i1; ++i) {
      const double xi = (double)xg[i];
      float a[8];
      load8(A + (long long)i * n + j0, a);
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] += (double)a[u] * xi;
    }
  } else {
    for (int u = 0; u < w; ++u) {
      const int col = c0 + u;
      const float *xg = x + gd.off_m + (long long)(col / n) * pm;
      const int j = col % n;
      for (int i = i0; i < i1; ++i) acc[u] += (double)a_val(A, (long long)i * n + j) * (double)xg[i];
    }
  }
  double *out = scratch + blockIdx.z * sstride + (long long)c * gd.n + c0;
  for (int u = 0; u < w; ++u) out[u] = acc[u];
}

__global__ __launch_bounds__(kBlock) void k_dense_bwd_reduce(
    const GroupDesc *G, const double *scratch, float *y, const InstScal *S,
    uint32_t mask, const float *sub, long long sstride) {
  const GroupDesc &gd = G[blockIdx.y];
  const int col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= gd.n) return;
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const int nch = min(kDenseChunks, (gd.plan.m + kBlock - 1) / kBlock);
  const double *sp = scratch + blockIdx.y * sstride + col;
  double acc = 0.0;
  for (int c = 0; c < nch; ++c) acc += sp[(long long)c * gd.n];
  const long long o = gd.off_n + col;
  float v = (float)acc;
  if (sub) v = v - sub[o];
  y[o] = v;
}

// The blocked A^T x of get_g (dense, W = 1): a thread owns C (4 or 8)
// consecutive columns and walks all rows in order (no split over rows), FP64
// within a block of rows, float across blocks.  y = g - sub.
template <class T, int U, int C>
__global__ __launch_bounds__(kBlock) void k_dense_bwd_blk(
    const GroupDesc *G, const float *x, float *y, const InstScal *S,
    uint32_t mask, const float *sub, int submat_mb) {
  const GroupDesc &gd = G[blockIdx.y];
  if (gd.dblk_S > 0) return;  // k_dense_bwd_bs / _bred handle this group
  const int n = gd.n;
  const int c0 = (blockIdx.x * kBlock + threadIdx.x) * C;
  if (c0 >= n) return;
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const int m = gd.m;
  const int rb = dense_block_rows(submat_mb, n);
  const float *xg = x + gd.off_m;
  const T *A = (const T *)gd.plan.A;
  const int w = min(C, n - c0);
  float gf[C];
#pragma unroll
  for (int u = 0; u < C; ++u) gf[u] = 0.0f;
  for (int y0 = 0; y0 < m; y0 += rb) {
    const int y1 = min(m, y0 + rb);
    double acc[C];
#pragma unroll
    for (int u = 0; u < C; ++u) acc[u] = 0.0;
    if (w == C && gd.plan.a_vec) {
      int i = y0;
      for (; i + U <= y1; i += U) {
        float a[U][C];
        double xi[U];
#pragma unroll
        for (int r = 0; r < U; ++r) {
          loadc<C>(A + (long long)(i + r) * n + c0, a[r]);
          xi[r] = (double)xg[i + r];
        }
#pragma unroll
        for (int r = 0; r < U; ++r)
#pragma unroll
          for (int u = 0; u < C; ++u) acc[u] = prod_add(acc[u], (double)a[r][u], xi[r]);
      }
      if (i < y1) {
        // the block's last rows in one round, their loads in flight together
        // (row by row they took a memory latency each: 11.5 -> 10.7 ms per
        // product of the 64-instance bf16 batch)
        float a[U][C];
        double xi[U];
#pragma unroll
        for (int r = 0; r < U; ++r) {
          if (i + r < y1) {
            loadc<C>(A + (long long)(i + r) * n + c0, a[r]);
            xi[r] = (double)xg[i + r];
          }
        }
#pragma unroll
        for (int r = 0; r < U; ++r)
          if (i + r < y1) {
#pragma unroll
            for (int u = 0; u < C; ++u) acc[u] = prod_add(acc[u], (double)a[r][u], xi[r]);
          }
      }
    } else {
      for (int i = y0; i < y1; ++i) {
        const double xi = (double)xg[i];
        for (int u = 0; u < w; ++u) acc[u] += (double)a_val(A, (long long)i * n + c0 + u) * xi;
      }
    }
#pragma unroll
    for (int u = 0; u < C; ++u) gf[u] = gf[u] + (float)acc[u];
  }
  const long long o = gd.off_n + c0;
  for (int u = 0; u < w; ++u) {
    float v = gf[u];
    if (sub) v = v - sub[o + u];
    y[o + u] = v;
  }
}

// Blocked A^T x with row slices (groups with few blocks): chunk c = b * S + s
// covers slice s of row block b; each thread sums 8 columns over the slice's
// rows in FP64 and writes the partials; k_dense_bwd_bred adds the S slices of
// a block in FP64, rounds once per block and adds the blocks in float.
constexpr int kDenseBlkChunks = 64;

template <class T, int U>
__global__ __launch_bounds__(kBlock) void k_dense_bwd_bs(
    const GroupDesc *G, const float *x, double *scratch, const InstScal *S,
    uint32_t mask) {
  const GroupDesc &gd = G[blockIdx.z];
  const int Sl = gd.dblk_S;
  if (Sl <= 0) return;
  const int n = gd.n, m = gd.m, rb = gd.dblk_rows;
  const int c0 = (blockIdx.x * kBlock + threadIdx.x) * 8;
  if (c0 >= n) return;
  if (mask && (S[blockIdx.z].flags & mask) != mask) return;
  const int nb = (m + rb - 1) / rb;
  const int c = blockIdx.y;
  if (c >= nb * Sl) return;
  const int b = c / Sl, sl = c % Sl;
  const int y0 = b * rb, y1 = min(m, y0 + rb);
  const int rps = (y1 - y0 + Sl - 1) / Sl;
  const int i0 = min(y1, y0 + sl * rps), i1 = min(y1, i0 + rps);
  const float *xg = x + gd.off_m;
  const T *A = (const T *)gd.plan.A;
  const int w = min(8, n - c0);
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (w == 8 && gd.plan.a_vec) {
    int i = i0;
    for (; i + U <= i1; i += U) {
      float a[U][8];
      double xi[U];
#pragma unroll
      for (int r = 0; r < U; ++r) {
        load8(A + (long long)(i + r) * n + c0, a[r]);
        xi[r] = (double)xg[i + r];
      }
#pragma unroll
      for (int r = 0; r < U; ++r)
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] = prod_add(acc[u], (double)a[r][u], xi[r]);
    }
    if (i < i1) {
      // the slice's last rows in one round
      float a[U][8];
      double xi[U];
#pragma unroll
      for (int r = 0; r < U; ++r) {
        if (i + r < i1) {
          load8(A + (long long)(i + r) * n + c0, a[r]);
          xi[r] = (double)xg[i + r];
        }
      }
#pragma unroll
      for (int r = 0; r < U; ++r)
        if (i + r < i1) {
#pragma unroll
          for (int u = 0; u < 8; ++u) acc[u] = prod_add(acc[u], (double)a[r][u], xi[r]);
        }
    }
  } else {
    for (int i = i0; i < i1; ++i) {
      const double xi = (double)xg[i];
      for (int u = 0; u < w; ++u) acc[u] += (double)a_val(A, (long long)i * n + c0 + u) * xi;
    }
  }
  double *out = scratch + gd.dblk_off + (long long)c * n + c0;
  for (int u = 0; u < w; ++u) out[u] = acc[u];
}

__global__ __launch_bounds__(kBlock) void k_dense_bwd_bred(
    const GroupDesc *G, const double *scratch, float *y, const InstScal *S,
    uint32_t mask, const float *sub) {
  const GroupDesc &gd = G[blockIdx.y];
  const int Sl = gd.dblk_S;
  if (Sl <= 0) return;
  const int col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= gd.n) return;
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const int nb = (gd.m + gd.dblk_rows - 1) / gd.dblk_rows;
  const double *sp = scratch + gd.dblk_off + col;
  float gf = 0.0f;
  for (int b = 0; b < nb; ++b) {
    double t = 0.0;
    for (int s2 = 0; s2 < Sl; ++s2) t += sp[(long long)(b * Sl + s2) * gd.n];
    gf = gf + (float)t;
  }
  const long long o = gd.off_n + col;
  y[o] = sub ? gf - sub[o] : gf;
}

// ---------------------------------------------------------------------------
// Multi-layer dense products on the matrix cores (FP64 MFMA)
// ---------------------------------------------------------------------------
// The reference's multi-layer mode applies one A to L images
// (2d/mat_mult.h:22-35: cblas_gemm over the layers).  Here both products are
// GEMMs over the layers on v_mfma_f64_16x16x4f64: float and bf16 entries and
// float vector values widen to FP64 exactly, so every product term is exact
// and the sums accumulate in FP64, rounded once to float as in the W = 1
// kernels (a different FP64 summation order, as between those kernels and
// the oracle).  A is read once per 16 layers instead of once per layer.
//
// Operand layout of v_mfma_f64_16x16x4f64 (measured on gfx950): lane l
// passes A-operand entry (row l % 16, k l / 16) and B-operand entry (k l / 16,
// column l % 16); result (i, j) is register i / 4 of lane 16 (i % 4) + j.
// A workgroup of NW wavefronts computes one 16 x 16 output tile; wave w sums
// its share of K (whole 16-wide K blocks), and wave 0 adds the NW partial
// tiles in wave order through LDS.
typedef double tv3_d4 __attribute__((ext_vector_type(4)));

// wavefronts per output tile group (TV3_MFMA_WAVES = 4, 8 or 16
// overrides).  8 measured best for one-tile workgroups on 64 x 64 layers
// sharing a 1229 x 4096 A: A X / A^T Y took
// 23.0 / 15.2 us at 8 layers and 38.6 / 42.9 us at 64 (4 waves: 31.8 / 15.9
// and 43.2 / 41.9; 16 waves: 21.8 / 15.7 and 43.2 / 58.8).  The kernels
// load one value per lane per MFMA (4 flops per loaded byte), so L2-to-CU
// bandwidth, not the matrix cores, bounds them.
int mfma_waves() {
  static const int w = [] {
    const char *e = std::getenv("TV3_MFMA_WAVES");
    const int v = e ? std::atoi(e) : 8;
    return (v == 4 || v == 16) ? v : 8;
  }();
  return w;
}

__device__ __forceinline__ float mf_load(const float *p) { return *p; }
__device__ __forceinline__ float mf_load(const __hip_bfloat16 *p) { return __bfloat162float(*p); }
// four consecutive entries, 16-byte (f32) / 8-byte (bf16) aligned
__device__ __forceinline__ void mf_load4(const float *p, float (&v)[4]) {
  const float4 q = *(const float4 *)p;
  v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
__device__ __forceinline__ void mf_load4(const __hip_bfloat16 *p, float (&v)[4]) { load4(p, v); }

// wave w's K range of NW: whole 16-wide blocks, an equal share each
template <int NW>
__device__ __forceinline__ void mf_krange(int K, int w, int &k0, int &k1) {
  const int nb = (K + 15) / 16, per = (nb + NW - 1) / NW;
  k0 = min(K, w * per * 16);
  k1 = min(K, (w + 1) * per * 16);
}

// adds the NW waves' partial tiles in wave order; true in wave 0
template <int NW>
__device__ __forceinline__ bool mf_reduce(tv3_d4 &acc) {
  __shared__ double red[NW - 1][4][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (w > 0)
#pragma unroll
    for (int r = 0; r < 4; ++r) red[w - 1][r][lane] = acc[r];
  __syncthreads();
  if (w > 0) return false;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double t = acc[r];
    for (int v = 0; v < NW - 1; ++v) t += red[v][r][lane];
    acc[r] = t;
  }
  return true;
}

// Each wave walks its K range in groups of kMfU 16-wide blocks: the next
// group's operands are loaded while the current group's 4 kMfU MFMAs run,
// and substep s of every block accumulates into acc[s], so four MFMA chains
// are in flight; the four accumulators are added in order at the end.
constexpr int kMfU = 4;

// loads one group of kMfU blocks starting at kb: the lane's four entries
// k = kb' + 4 q + s of every block kb' < k1 (zero beyond the operand's K and
// the wave's range).  FWD: both operands run along K contiguously; else the
// B operand is column `col` of a row-major A with row stride n.
template <class T, bool VEC, bool FWD>
__device__ __forceinline__ void mf_group(const T *Ab, const float *xl, bool aok, bool xok, int kb,
                                         int k1, int K, int q, long long n, float (&av)[kMfU][4],
                                         float (&xv)[kMfU][4]) {
#pragma unroll
  for (int u = 0; u < kMfU; ++u) {
    const int k = kb + 16 * u + 4 * q;
#pragma unroll
    for (int s = 0; s < 4; ++s) { av[u][s] = 0.0f; xv[u][s] = 0.0f; }
    if (kb + 16 * u >= k1) continue;
    if (VEC && k + 4 <= K) {
      if (xok) mf_load4(xl + k, xv[u]);
      if (FWD && aok) mf_load4(Ab + k, av[u]);
    } else {
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (k + s < K) {
          if (xok) xv[u][s] = xl[k + s];
          if (FWD && aok) av[u][s] = mf_load(Ab + k + s);
        }
    }
    if (!FWD)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        if (aok && k + s < K) av[u][s] = mf_load(Ab + (long long)(k + s) * n);
  }
}

// the four-accumulator MFMA walk of one wave's K range [k0, k1)
template <class T, bool VEC, bool FWD>
__device__ __forceinline__ tv3_d4 mf_walk(const T *Ab, const float *xl, bool aok, bool xok, int k0,
                                          int k1, int K, int q, long long n) {
  tv3_d4 acc[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) acc[s] = tv3_d4{0.0, 0.0, 0.0, 0.0};
  float av[kMfU][4], xv[kMfU][4];
  mf_group<T, VEC, FWD>(Ab, xl, aok, xok, k0, k1, K, q, n, av, xv);
  for (int kb = k0; kb < k1; kb += 16 * kMfU) {
    float an[kMfU][4], xn[kMfU][4];
    const bool more = kb + 16 * kMfU < k1;
    if (more) mf_group<T, VEC, FWD>(Ab, xl, aok, xok, kb + 16 * kMfU, k1, K, q, n, an, xn);
#pragma unroll
    for (int u = 0; u < kMfU; ++u)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)xv[u][s], (double)av[u][s], acc[s], 0, 0, 0);
    if (more) {
#pragma unroll
      for (int u = 0; u < kMfU; ++u)
#pragma unroll
        for (int s = 0; s < 4; ++s) { av[u][s] = an[u][s]; xv[u][s] = xn[u][s]; }
    }
  }
  tv3_d4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = ((acc[0][i] + acc[1][i]) + acc[2][i]) + acc[3][i];
  return r;
}

// Y (L x m) = X (L x n) A^T: layer l's A x at y[off_m + l m + i].  Grid:
// (row tiles of 16, layer tiles of 16, groups); VEC: A rows and the layers'
// n-vectors take 16-byte vector loads (a_vec).
template <class T, bool VEC, int NW>
__global__ __launch_bounds__(64 * NW) void k_dense_fwd_mfma(const GroupDesc *G, const float *x,
                                                            float *y, const InstScal *S,
                                                            uint32_t mask) {
  const GroupDesc &gd = G[blockIdx.z];
  const int n = gd.plan.n, pm = gd.plan.m, L = gd.layers;
  const int i0 = blockIdx.x * 16, l0 = blockIdx.y * 16;
  if (i0 >= pm || l0 >= L) return;
  if (mask && (S[blockIdx.z].flags & mask) != mask) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = lane & 15, q = lane >> 4;
  const int row = i0 + c, lyr = l0 + c;
  const bool rok = row < pm, lok = lyr < L;
  const T *Ar = (const T *)gd.plan.A + (long long)(rok ? row : 0) * n;
  const float *xl = x + gd.off_n + (long long)(lok ? lyr : 0) * n;
  int k0, k1;
  mf_krange<NW>(n, w, k0, k1);
  tv3_d4 acc = mf_walk<T, VEC, true>(Ar, xl, rok, lok, k0, k1, n, q, n);
  if (!mf_reduce<NW>(acc)) return;
  // register r of this lane: layer l0 + 4 r + q, row i0 + c
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int l = l0 + 4 * r + q;
    if (l < L && rok) y[gd.off_m + (long long)l * pm + row] = (float)acc[r];
  }
}

// G (L x n) = Y (L x m) A: layer l's A^T y at y[off_n + l n + j], minus sub.
// Grid: (column tiles of 16, layer tiles of 16, groups); VEC: the layers'
// m-vectors take 16-byte vector loads (m % 4 == 0).
template <class T, bool VEC, int NW>
__global__ __launch_bounds__(64 * NW) void k_dense_bwd_mfma(const GroupDesc *G, const float *x,
                                                            float *y, const InstScal *S,
                                                            uint32_t mask, const float *sub) {
  const GroupDesc &gd = G[blockIdx.z];
  const int n = gd.plan.n, pm = gd.plan.m, L = gd.layers;
  const int j0 = blockIdx.x * 16, l0 = blockIdx.y * 16;
  if (j0 >= n || l0 >= L) return;
  if (mask && (S[blockIdx.z].flags & mask) != mask) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = lane & 15, q = lane >> 4;
  const int col = j0 + c, lyr = l0 + c;
  const bool cok = col < n, lok = lyr < L;
  const T *Ac = (const T *)gd.plan.A + (cok ? col : 0);
  const float *xl = x + gd.off_m + (long long)(lok ? lyr : 0) * pm;
  int k0, k1;
  mf_krange<NW>(pm, w, k0, k1);
  tv3_d4 acc = mf_walk<T, VEC, false>(Ac, xl, cok, lok, k0, k1, pm, q, n);
  if (!mf_reduce<NW>(acc)) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int l = l0 + 4 * r + q;
    if (l < L && cok) {
      const long long o = gd.off_n + (long long)l * n + col;
      float v = (float)acc[r];
      if (sub) v = v - sub[o];
      y[o] = v;
    }
  }
}

// Register-blocked variants: a workgroup computes TL x TR output tiles (TL
// layer tiles by TR row or column tiles), so each operand a lane loads feeds
// TR (layer operands) or TL (A operands) MFMAs instead of one.  Every tile
// is summed exactly as in the one-tile kernels above (the same wave split of
// K, the same four accumulator chains per 16-wide block, the same final and
// wave-order additions), so both give the same bits.
constexpr int kMfUb = 2;  // 16-wide blocks per load group

template <class T, bool VEC, bool FWD, int TL, int TR>
__device__ __forceinline__ void mfb_group(const T *const (&Ab)[TR], const float *const (&xl)[TL],
                                          const bool (&aok)[TR], const bool (&xok)[TL], int kb, int k1,
                                          int K, int q, long long n, float (&av)[kMfUb][TR][4],
                                          float (&xv)[kMfUb][TL][4]) {
#pragma unroll
  for (int u = 0; u < kMfUb; ++u) {
    const int k = kb + 16 * u + 4 * q;
    const bool live = kb + 16 * u < k1;
#pragma unroll
    for (int t = 0; t < TL; ++t) {
#pragma unroll
      for (int s = 0; s < 4; ++s) xv[u][t][s] = 0.0f;
      if (!live || !xok[t]) continue;
      if (VEC && k + 4 <= K) {
        mf_load4(xl[t] + k, xv[u][t]);
      } else {
#pragma unroll
        for (int s = 0; s < 4; ++s)
          if (k + s < K) xv[u][t][s] = xl[t][k + s];
      }
    }
#pragma unroll
    for (int t = 0; t < TR; ++t) {
#pragma unroll
      for (int s = 0; s < 4; ++s) av[u][t][s] = 0.0f;
      if (!live || !aok[t]) continue;
      if (FWD) {
        if (VEC && k + 4 <= K) {
          mf_load4(Ab[t] + k, av[u][t]);
        } else {
#pragma unroll
          for (int s = 0; s < 4; ++s)
            if (k + s < K) av[u][t][s] = mf_load(Ab[t] + k + s);
        }
      } else {
#pragma unroll
        for (int s = 0; s < 4; ++s)
          if (k + s < K) av[u][t][s] = mf_load(Ab[t] + (long long)(k + s) * n);
      }
    }
  }
}

template <class T, bool VEC, bool FWD, int TL, int TR>
__device__ __forceinline__ void mfb_walk(const T *const (&Ab)[TR], const float *const (&xl)[TL],
                                         const bool (&aok)[TR], const bool (&xok)[TL], int k0, int k1,
                                         int K, int q, long long n, tv3_d4 (&out)[TL][TR]) {
  tv3_d4 acc[TL][TR][4];
#pragma unroll
  for (int a = 0; a < TL; ++a)
#pragma unroll
    for (int b = 0; b < TR; ++b)
#pragma unroll
      for (int s = 0; s < 4; ++s) acc[a][b][s] = tv3_d4{0.0, 0.0, 0.0, 0.0};
  float av[kMfUb][TR][4], xv[kMfUb][TL][4];
  mfb_group<T, VEC, FWD, TL, TR>(Ab, xl, aok, xok, k0, k1, K, q, n, av, xv);
  for (int kb = k0; kb < k1; kb += 16 * kMfUb) {
    float an[kMfUb][TR][4], xn[kMfUb][TL][4];
    const bool more = kb + 16 * kMfUb < k1;
    if (more) mfb_group<T, VEC, FWD, TL, TR>(Ab, xl, aok, xok, kb + 16 * kMfUb, k1, K, q, n, an, xn);
#pragma unroll
    for (int u = 0; u < kMfUb; ++u)
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int a = 0; a < TL; ++a)
#pragma unroll
          for (int b = 0; b < TR; ++b)
            acc[a][b][s] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)xv[u][a][s], (double)av[u][b][s],
                                                                acc[a][b][s], 0, 0, 0);
    if (more) {
#pragma unroll
      for (int u = 0; u < kMfUb; ++u) {
#pragma unroll
        for (int a = 0; a < TL; ++a)
#pragma unroll
          for (int s = 0; s < 4; ++s) xv[u][a][s] = xn[u][a][s];
#pragma unroll
        for (int b = 0; b < TR; ++b)
#pragma unroll
          for (int s = 0; s < 4; ++s) av[u][b][s] = an[u][b][s];
      }
    }
  }
#pragma unroll
  for (int a = 0; a < TL; ++a)
#pragma unroll
    for (int b = 0; b < TR; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        out[a][b][i] = ((acc[a][b][0][i] + acc[a][b][1][i]) + acc[a][b][2][i]) + acc[a][b][3][i];
}

// adds the NW waves' partial tiles in wave order (tile by tile, as
// mf_reduce); true in wave 0
template <int NW, int NT>
__device__ __forceinline__ bool mfb_reduce(tv3_d4 *acc) {
  __shared__ double red[NW - 1][NT][4][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (w > 0)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[w - 1][t][r][lane] = acc[t][r];
  __syncthreads();
  if (w > 0) return false;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double v = acc[t][r];
      for (int k = 0; k < NW - 1; ++k) v += red[k][t][r][lane];
      acc[t][r] = v;
    }
  return true;
}

// Y (L x m) = X (L x n) A^T over TL x TR tiles of 16 layers x 16 rows.
// Grid: (row tile groups, layer tile groups, groups)
template <class T, bool VEC, int NW, int TL, int TR>
__global__ __launch_bounds__(64 * NW) void k_dense_fwd_mfmab(const GroupDesc *G, const float *x,
                                                             float *y, const InstScal *S,
                                                             uint32_t mask) {
  const GroupDesc &gd = G[blockIdx.z];
  const int n = gd.plan.n, pm = gd.plan.m, L = gd.layers;
  const int i0 = blockIdx.x * 16 * TR, l0 = blockIdx.y * 16 * TL;
  if (i0 >= pm || l0 >= L) return;
  if (mask && (S[blockIdx.z].flags & mask) != mask) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = lane & 15, q = lane >> 4;
  const T *Ar[TR];
  bool rok[TR];
#pragma unroll
  for (int b = 0; b < TR; ++b) {
    const int row = i0 + 16 * b + c;
    rok[b] = row < pm;
    Ar[b] = (const T *)gd.plan.A + (long long)(rok[b] ? row : 0) * n;
  }
  const float *xl[TL];
  bool lok[TL];
#pragma unroll
  for (int a = 0; a < TL; ++a) {
    const int lyr = l0 + 16 * a + c;
    lok[a] = lyr < L;
    xl[a] = x + gd.off_n + (long long)(lok[a] ? lyr : 0) * n;
  }
  int k0, k1;
  mf_krange<NW>(n, w, k0, k1);
  tv3_d4 acc[TL][TR];
  mfb_walk<T, VEC, true, TL, TR>(Ar, xl, rok, lok, k0, k1, n, q, n, acc);
  if (!mfb_reduce<NW, TL * TR>(&acc[0][0])) return;
#pragma unroll
  for (int a = 0; a < TL; ++a)
#pragma unroll
    for (int b = 0; b < TR; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int l = l0 + 16 * a + 4 * r + q, row = i0 + 16 * b + c;
        if (l < L && row < pm) y[gd.off_m + (long long)l * pm + row] = (float)acc[a][b][r];
      }
}

// G (L x n) = Y (L x m) A over TL x TR tiles of 16 layers x 16 columns,
// minus sub.  Grid: (column tile groups, layer tile groups, groups)
template <class T, bool VEC, int NW, int TL, int TR>
__global__ __launch_bounds__(64 * NW) void k_dense_bwd_mfmab(const GroupDesc *G, const float *x,
                                                             float *y, const InstScal *S,
                                                             uint32_t mask, const float *sub) {
  const GroupDesc &gd = G[blockIdx.z];
  const int n = gd.plan.n, pm = gd.plan.m, L = gd.layers;
  const int j0 = blockIdx.x * 16 * TR, l0 = blockIdx.y * 16 * TL;
  if (j0 >= n || l0 >= L) return;
  if (mask && (S[blockIdx.z].flags & mask) != mask) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = lane & 15, q = lane >> 4;
  const T *Ac[TR];
  bool cok[TR];
#pragma unroll
  for (int b = 0; b < TR; ++b) {
    const int col = j0 + 16 * b + c;
    cok[b] = col < n;
    Ac[b] = (const T *)gd.plan.A + (cok[b] ? col : 0);
  }
  const float *xl[TL];
  bool lok[TL];
#pragma unroll
  for (int a = 0; a < TL; ++a) {
    const int lyr = l0 + 16 * a + c;
    lok[a] = lyr < L;
    xl[a] = x + gd.off_m + (long long)(lok[a] ? lyr : 0) * pm;
  }
  int k0, k1;
  mf_krange<NW>(pm, w, k0, k1);
  tv3_d4 acc[TL][TR];
  mfb_walk<T, VEC, false, TL, TR>(Ac, xl, cok, lok, k0, k1, pm, q, n, acc);
  if (!mfb_reduce<NW, TL * TR>(&acc[0][0])) return;
#pragma unroll
  for (int a = 0; a < TL; ++a)
#pragma unroll
    for (int b = 0; b < TR; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int l = l0 + 16 * a + 4 * r + q, col = j0 + 16 * b + c;
        if (l < L && col < n) {
          const long long o = gd.off_n + (long long)l * n + col;
          float v = (float)acc[a][b][r];
          if (sub) v = v - sub[o];
          y[o] = v;
        }
      }
}

// output tiles per workgroup of the multi-layer GEMMs, for A X (f) and
// A^T Y (b): 1 one 16 x 16 tile (k_dense_*_mfma), 2 two by two tiles, 3 (A X
// only) two layer tiles by one row tile (k_dense_*_mfmab); TV3_MFMA_BLOCK=f,b
// overrides.  Measured on 64 layers of 64 x 64 sharing a 1229 x 4096 A
// (eight wavefronts per workgroup): A^T Y 42.4 -> 34.5 us with 2 x 2 tiles
// (256 workgroups); A X 38.8 -> 57.7 us with 2 x 2 (78 workgroups: too few
// to fill the chip) and 36.8 us with 2 x 1; sixteen wavefronts per
// workgroup leave the 2 x 2 kernels 128 registers and spill them.  (Read
// at every launch, so tests can switch it; the kernels give the same bits.)
int mfma_block(bool fwd) {
  int f = 3, b = 2;
  if (const char *e = std::getenv("TV3_MFMA_BLOCK")) (void)std::sscanf(e, "%d,%d", &f, &b);
  return fwd ? (f == 2 || f == 3 ? f : 1) : (b == 1 ? 1 : 2);
}

template <>
void launch_fwd<1>(hipStream_t s, const GroupDesc *G_dev, int G, int max_m,
                   int kind, int dtype, const float *x, float *y,
                   const InstScal *S, uint32_t mask) {
  if (kind == kDense && dense_tune_cur().mfma) {
    // tiles of the largest A and layer count; smaller groups exit early
    const DenseTune t = dense_tune_cur();
    const int tb = mfma_block(true);
    const int tl = tb == 1 ? 1 : 2, tr = tb == 2 ? 2 : 1;
    const dim3 g((t.mfma_m + 16 * tr - 1) / (16 * tr), (t.mfma_layers + 16 * tl - 1) / (16 * tl), G);
#define TV3_MF(TT, VV, NW)                                                                                    \
  if (tb == 2)                                                                                              \
    hipLaunchKernelGGL((k_dense_fwd_mfmab<TT, VV, NW, 2, 2>), g, dim3(64 * NW), 0, s, G_dev, x, y, S, mask); \
  else if (tb == 3)                                                                                         \
    hipLaunchKernelGGL((k_dense_fwd_mfmab<TT, VV, NW, 2, 1>), g, dim3(64 * NW), 0, s, G_dev, x, y, S, mask); \
  else                                                                                                      \
    hipLaunchKernelGGL((k_dense_fwd_mfma<TT, VV, NW>), g, dim3(64 * NW), 0, s, G_dev, x, y, S, mask)
#define TV3_MFW(TT, VV)                           \
  if (mfma_waves() == 4) TV3_MF(TT, VV, 4);       \
  else if (mfma_waves() == 16) TV3_MF(TT, VV, 16); \
  else TV3_MF(TT, VV, 8);
    if (dtype == TV3_BF16) {
      if (t.mfma_vec_f) { TV3_MFW(__hip_bfloat16, true) } else { TV3_MFW(__hip_bfloat16, false) }
    } else {
      if (t.mfma_vec_f) { TV3_MFW(float, true) } else { TV3_MFW(float, false) }
    }
#undef TV3_MFW
#undef TV3_MF
    (void)max_m;
    return;
  }
  if (kind == kDense) {
    const int R = dense_tune_cur().rows;
    dim3 g((max_m + 4 * R - 1) / (4 * R), G);
#define TV3_DFWD(TT, RR) hipLaunchKernelGGL((k_dense_fwd_w1<TT, RR>), g, dim3(kBlock), 0, s, G_dev, x, y, S, mask)
    if (dtype == TV3_BF16) {
      if (R == 4) TV3_DFWD(__hip_bfloat16, 4);
      else if (R == 2) TV3_DFWD(__hip_bfloat16, 2);
      else TV3_DFWD(__hip_bfloat16, 1);
    } else {
      if (R == 4) TV3_DFWD(float, 4);
      else if (R == 2) TV3_DFWD(float, 2);
      else TV3_DFWD(float, 1);
    }
#undef TV3_DFWD
    return;
  }
  dim3 grid((max_m + kBlock / kSubF - 1) / (kBlock / kSubF), G);
  hipLaunchKernelGGL(k_fwd_w1<kSubF>, grid, dim3(kBlock), 0, s, G_dev, x, y, S, mask);
}

template <>
void launch_fwd<64>(hipStream_t s, const GroupDesc *G_dev, int G, int max_m,
                    int kind, int dtype, const float *x, float *y,
                    const InstScal *S, uint32_t mask) {
  (void)dtype;
  dim3 grid((max_m + 3) / 4, G);
  if (kind == kGeometry && proj_vec()) {
    hipLaunchKernelGGL(k_fwd_w64q, grid, dim3(kBlock), 0, s, G_dev, x, y, S, mask);
  } else {
    hipLaunchKernelGGL(k_fwd_w64, grid, dim3(kBlock), 0, s, G_dev, x, y, S, mask);
  }
}

template <>
void launch_bwd<1>(hipStream_t s, const GroupDesc *G_dev, int G, int max_n,
                   int kind, int dtype, const float *x, float *y,
                   const InstScal *S, uint32_t mask, const float *sub,
                   float *scratch, int submat_mb) {
  if (kind == kDense && dense_tune_cur().mfma) {
    const DenseTune t = dense_tune_cur();
    const int tb = mfma_block(false);
    const dim3 g((t.mfma_n + 16 * tb - 1) / (16 * tb), (t.mfma_layers + 16 * tb - 1) / (16 * tb), G);
#define TV3_MB(TT, VV, NW)                                                                                     \
  if (tb == 2)                                                                                               \
    hipLaunchKernelGGL((k_dense_bwd_mfmab<TT, VV, NW, 2, 2>), g, dim3(64 * NW), 0, s, G_dev, x, y, S, mask, \
                       sub);                                                                                 \
  else                                                                                                       \
    hipLaunchKernelGGL((k_dense_bwd_mfma<TT, VV, NW>), g, dim3(64 * NW), 0, s, G_dev, x, y, S, mask, sub)
#define TV3_MBW(TT, VV)                           \
  if (mfma_waves() == 4) TV3_MB(TT, VV, 4);       \
  else if (mfma_waves() == 16) TV3_MB(TT, VV, 16); \
  else TV3_MB(TT, VV, 8);
    if (dtype == TV3_BF16) {
      if (t.mfma_vec_b) { TV3_MBW(__hip_bfloat16, true) } else { TV3_MBW(__hip_bfloat16, false) }
    } else {
      if (t.mfma_vec_b) { TV3_MBW(float, true) } else { TV3_MBW(float, false) }
    }
#undef TV3_MBW
#undef TV3_MB
    (void)max_n;
    (void)scratch;
    (void)submat_mb;
    return;
  }
  if (kind == kDense && submat_mb > 0) {
    const DenseTune tn = dense_tune_cur();
    const int Ur = tn.unroll, Cc = tn.cols;
    dim3 g((max_n + Cc * kBlock - 1) / (Cc * kBlock), G);
#define TV3_DBLK(TT, UU, CC) hipLaunchKernelGGL((k_dense_bwd_blk<TT, UU, CC>), g, dim3(kBlock), 0, s, G_dev, x, y, \
                                               S, mask, sub, submat_mb)
#define TV3_DBLK_U(TT, CC)              \
  if (Ur == 8) TV3_DBLK(TT, 8, CC);     \
  else if (Ur == 4) TV3_DBLK(TT, 4, CC); \
  else if (Ur == 2) TV3_DBLK(TT, 2, CC); \
  else TV3_DBLK(TT, 1, CC);
    if (tn.no_blk) {
      // no group of this batch takes the whole-row kernel
    } else if (dtype == TV3_BF16) {
      if (Cc == 2) { TV3_DBLK_U(__hip_bfloat16, 2) } else if (Cc == 4) { TV3_DBLK_U(__hip_bfloat16, 4) } else { TV3_DBLK_U(__hip_bfloat16, 8) }
    } else {
      if (Cc == 2) { TV3_DBLK_U(float, 2) } else if (Cc == 4) { TV3_DBLK_U(float, 4) } else { TV3_DBLK_U(float, 8) }
    }
#undef TV3_DBLK_U
#undef TV3_DBLK
    if (tn.no_slice) return;
    // groups in slice mode
    dim3 g1((max_n + 8 * kBlock - 1) / (8 * kBlock), kDenseBlkChunks, G);
#define TV3_DBS(TT, UU) hipLaunchKernelGGL((k_dense_bwd_bs<TT, UU>), g1, dim3(kBlock), 0, s, G_dev, x, \
                                          (double *)scratch, S, mask)
    if (dtype == TV3_BF16) {
      if (Ur >= 4) TV3_DBS(__hip_bfloat16, 4);
      else if (Ur == 2) TV3_DBS(__hip_bfloat16, 2);
      else TV3_DBS(__hip_bfloat16, 1);
    } else {
      if (Ur >= 4) TV3_DBS(float, 4);
      else if (Ur == 2) TV3_DBS(float, 2);
      else TV3_DBS(float, 1);
    }
#undef TV3_DBS
    dim3 g2((max_n + kBlock - 1) / kBlock, G);
    hipLaunchKernelGGL(k_dense_bwd_bred, g2, dim3(kBlock), 0, s, G_dev, (const double *)scratch, y, S,
                       mask, sub);
    return;
  }
  if (kind == kDense) {
    // scratch holds kDenseChunks * max_n doubles per group
    const long long sstride = (long long)kDenseChunks * max_n;
    dim3 g1((max_n + 8 * kBlock - 1) / (8 * kBlock), kDenseChunks, G);
    const int Ur = dense_tune_cur().unroll;
#define TV3_DBWD(TT, UU) hipLaunchKernelGGL((k_dense_bwd_part<TT, UU>), g1, dim3(kBlock), 0, s, G_dev, x, \
                                           (double *)scratch, S, mask, sstride)
    if (dtype == TV3_BF16) {
      if (Ur >= 4) TV3_DBWD(__hip_bfloat16, 4);
      else if (Ur == 2) TV3_DBWD(__hip_bfloat16, 2);
      else TV3_DBWD(__hip_bfloat16, 1);
    } else {
      if (Ur >= 4) TV3_DBWD(float, 4);
      else if (Ur == 2) TV3_DBWD(float, 2);
      else TV3_DBWD(float, 1);
    }
#undef TV3_DBWD
    dim3 g2((max_n + kBlock - 1) / kBlock, G);
    hipLaunchKernelGGL(k_dense_bwd_reduce, g2, dim3(kBlock), 0, s, G_dev,
                       (const double *)scratch, y, S, mask, sub, sstride);
    return;
  }
  if (kind == kGeometry && stage_enabled()) {
    dim3 grid((max_n + kBlock - 1) / kBlock + 1 + kHeavyBlocks, G);
    const int ni = stage_depth();
#define TV3_W1S(NI_, NR_) hipLaunchKernelGGL((k_bwd_w1s<kSubT, NI_, NR_>), grid, dim3(kBlock), 0, s, G_dev, x, y, S, mask, sub)
    if (stage_rays() <= 6144) {
      if (ni <= 8) TV3_W1S(8, 6144); else if (ni <= 12) TV3_W1S(12, 6144); else TV3_W1S(16, 6144);
    } else {
      if (ni <= 8) TV3_W1S(8, 8192); else if (ni <= 12) TV3_W1S(12, 8192); else TV3_W1S(16, 8192);
    }
#undef TV3_W1S
    return;
  }
  // one block more than the columns need: the voxels no ray crosses take
  // blocks of their own (k_bwd_w1)
  dim3 grid((max_n + kBlock / kSubT - 1) / (kBlock / kSubT) + 1, G);
  hipLaunchKernelGGL(k_bwd_w1<kSubT>, grid, dim3(kBlock), 0, s, G_dev, x, y, S, mask,
                     sub);
}

template <>
void launch_bwd<64>(hipStream_t s, const GroupDesc *G_dev, int G, int max_n,
                    int kind, int dtype, const float *x, float *y,
                    const InstScal *S, uint32_t mask, const float *sub,
                    float *scratch, int submat_mb) {
  (void)dtype;
  (void)scratch;
  if (kind == kGeometry && proj_vec()) {
    const int per = (kBlock / 64) * kBwdCols;
    hipLaunchKernelGGL(k_bwd_w64q, dim3((max_n + per - 1) / per, G), dim3(kBlock), 0, s, G_dev,
                       x, y, S, mask, sub);
  } else {
    dim3 grid((max_n + 3) / 4, G);
    hipLaunchKernelGGL(k_bwd_w64, grid, dim3(kBlock), 0, s, G_dev, x, y, S, mask, sub,
                       kind == kDense ? submat_mb : 0);
  }
}

template <>
void launch_fwd<16>(hipStream_t s, const GroupDesc *G_dev, int G, int max_m,
                    int kind, int dtype, const float *x, float *y,
                    const InstScal *S, uint32_t mask) {
  (void)kind;
  (void)dtype;
  const int RB = (max_m + 3) / 4;
  if (G >= kXcd) {
    const int nb = kXcd * ((G + kXcd - 1) / kXcd) * RB;
    hipLaunchKernelGGL(k_fwd_w16, dim3(nb), dim3(kBlock), 0, s, G_dev, G, RB, 1,
                       x, y, S, mask);
  } else {
    hipLaunchKernelGGL(k_fwd_w16, dim3(RB, G), dim3(kBlock), 0, s, G_dev, G, RB,
                       0, x, y, S, mask);
  }
}

template <>
void launch_bwd<16>(hipStream_t s, const GroupDesc *G_dev, int G, int max_n,
                    int kind, int dtype, const float *x, float *y,
                    const InstScal *S, uint32_t mask, const float *sub,
                    float *scratch, int submat_mb) {
  (void)kind;
  (void)dtype;
  (void)scratch;
  const int RB = (max_n + 3) / 4;
  if (G >= kXcd) {
    const int nb = kXcd * ((G + kXcd - 1) / kXcd) * RB;
    hipLaunchKernelGGL(k_bwd_w16, dim3(nb), dim3(kBlock), 0, s, G_dev, G, RB, 1,
                       x, y, S, mask, sub);
  } else {
    hipLaunchKernelGGL(k_bwd_w16, dim3(RB, G), dim3(kBlock), 0, s, G_dev, G, RB,
                       0, x, y, S, mask, sub);
  }
}

// ---------------------------------------------------------------------------
// FP64 vectors, W = 1: the reference's overloads with T_comp = double
// (mat<double>, sparse_mat<double> and geometry with mat<double> U,
// 3d/tval3_cpu_3d.cpp:47, 88, 124; 2d/tval3_cpu_2d.cpp:26-29).  Entries,
// vector values and every product are double (a geometry entry is
// x * (double)scale, mult_elementN with T = double, 3d/mat_mult.h:121-139);
// the sums are FP64 like the reference's but run in a different order (lane
// sums, row chunks), so results agree to FP64 rounding.  Layers as in the
// float W = 1 kernels.
// ---------------------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(kBlock) void k_fwd_w1d(const GroupDesc *G,
                                                    const double *x, double *y,
                                                    const InstScal *S,
                                                    uint32_t mask) {
  const GroupDesc &gd = G[blockIdx.y];
  const int sl = threadIdx.x & (L - 1);
  const int row = blockIdx.x * (kBlock / L) + (threadIdx.x / L);
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const PlanDev &P = gd.plan;
  const int ar = row % P.m;
  const double *xg = x + gd.off_n + (long long)(row / P.m) * P.n;
  double acc = 0.0;
  if (row < gd.m) {
    constexpr int U = kW1Unroll;
    if (P.kind == kGeometry) {
      const RayDesc rd = P.rays[ar];
      const double sc = (double)P.scale;
      for (int j = sl; j <= rd.L; j += L * U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int jj = j + u * L;
          v[u] = jj <= rd.L ? xg[ray_lin(rd, jj)] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (j + u * L <= rd.L) acc += v[u] * sc;
      }
    } else if (P.kind == kSparse) {
      const int k0 = P.r_ptr[ar], k1 = P.r_ptr[ar + 1];
      for (int k = k0 + sl; k < k1; k += L * U) {
        int id[U];
        double w[U], v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int kk = k + u * L;
          id[u] = kk < k1 ? P.r_ind[kk] : -1;
          w[u] = kk < k1 ? P.r_vald[kk] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = id[u] >= 0 ? xg[id[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (id[u] >= 0) acc += w[u] * v[u];
      }
    } else {
      const double *A = (const double *)P.A + (long long)ar * P.n;
      for (int j = sl; j < P.n; j += L) acc += A[j] * xg[j];
    }
  }
  acc = lane_sum<L>(acc);
  if (sl == 0 && row < gd.m) y[gd.off_m + (long long)(row / P.m) * P.m + ar] = acc;
}

template <int L>
__global__ __launch_bounds__(kBlock) void k_bwd_w1d(const GroupDesc *G,
                                                    const double *x, double *y,
                                                    const InstScal *S,
                                                    uint32_t mask,
                                                    const double *sub) {
  const GroupDesc &gd = G[blockIdx.y];
  const int sl = threadIdx.x & (L - 1);
  const int col = blockIdx.x * (kBlock / L) + (threadIdx.x / L);
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const PlanDev &P = gd.plan;
  const bool geo = P.kind == kGeometry;
  const bool tiled = geo && P.t_col;
  int ac = col % P.n;
  const double *xg = x + gd.off_m + (long long)(col / P.n) * P.m;
  double acc = 0.0;
  if (col < gd.n) {
    constexpr int U = kW1Unroll;
    const int *ptr = tiled ? P.t_ptr : geo ? P.g_cptr : P.c_ptr;
    const int *ind = tiled ? P.t_ray : geo ? P.g_cray : P.c_ind;
    const int k0 = ptr[ac], k1 = ptr[ac + 1];
    if (tiled) ac = P.t_col[col];
    const double sc = (double)P.scale;
    for (int k = k0 + sl; k < k1; k += L * U) {
      int id[U];
      double w[U], v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int kk = k + u * L;
        id[u] = kk < k1 ? ind[kk] : -1;
        w[u] = (!geo && kk < k1) ? P.c_vald[kk] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = id[u] >= 0 ? xg[id[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (id[u] >= 0) acc += geo ? v[u] * sc : w[u] * v[u];
    }
  }
  acc = lane_sum<L>(acc);
  if (sl == 0 && col < gd.n) {
    const long long o = gd.off_n + (long long)(col / P.n) * P.n + ac;
    y[o] = sub ? acc - sub[o] : acc;
  }
}

// dense A^T x: chunk c of a group covers rows [c*R, min(m, (c+1)*R)); a
// thread owns one column (a wavefront reads 512 contiguous bytes of a row)
// and writes its FP64 chunk sum; k_dense_bwd_reduce_d adds the chunks.
__global__ __launch_bounds__(kBlock) void k_dense_bwd_part_d(
    const GroupDesc *G, const double *x, double *scratch, const InstScal *S,
    uint32_t mask, long long sstride) {
  const GroupDesc &gd = G[blockIdx.z];
  const int n = gd.plan.n, pm = gd.plan.m;
  const int col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= gd.n) return;
  if (mask && (S[blockIdx.z].flags & mask) != mask) return;
  const int nch = min(kDenseChunks, (pm + 63) / 64);
  const int c = blockIdx.y;
  if (c >= nch) return;
  const int R = (pm + nch - 1) / nch;
  const int i0 = c * R, i1 = min(pm, i0 + R);
  const double *A = (const double *)gd.plan.A + col % n;
  const double *xg = x + gd.off_m + (long long)(col / n) * pm;
  double acc = 0.0;
  int i = i0;
  for (; i + 4 <= i1; i += 4) {
    double a[4], xi[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a[r] = A[(long long)(i + r) * n];
      xi[r] = xg[i + r];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) acc += a[r] * xi[r];
  }
  for (; i < i1; ++i) acc += A[(long long)i * n] * xg[i];
  scratch[blockIdx.z * sstride + (long long)c * gd.n + col] = acc;
}

__global__ __launch_bounds__(kBlock) void k_dense_bwd_reduce_d(
    const GroupDesc *G, const double *scratch, double *y, const InstScal *S,
    uint32_t mask, const double *sub, long long sstride) {
  const GroupDesc &gd = G[blockIdx.y];
  const int col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= gd.n) return;
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const int nch = min(kDenseChunks, (gd.plan.m + 63) / 64);
  const double *sp = scratch + blockIdx.y * sstride + col;
  double acc = 0.0;
  for (int c = 0; c < nch; ++c) acc += sp[(long long)c * gd.n];
  const long long o = gd.off_n + col;
  y[o] = sub ? acc - sub[o] : acc;
}

// Double vectors with a bf16-stored A: the FP64 solve of the bf16-rounded
// matrix (entries widen exactly).  y = A x: four rows per wavefront share
// the x loads; with vector rows (a_vec) lane l sums entries 8 l + 512 t + u
// (u = 0..7) in order, else entries l + 64 t; FP64 lane sums and a fixed
// butterfly, so each row's order is the same whether or not its tile
// crosses a layer boundary.
__global__ __launch_bounds__(kBlock) void k_dense_fwd_db(const GroupDesc *G, const double *x, double *y,
                                                         const InstScal *S, uint32_t mask) {
  const GroupDesc &gd = G[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * 4;
  if (row0 >= gd.m) return;
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const int n = gd.plan.n, pm = gd.plan.m;
  const int nr = min(4, gd.m - row0);
  const __hip_bfloat16 *A = (const __hip_bfloat16 *)gd.plan.A;
  const bool vec = gd.plan.a_vec != 0;
  auto row_sum = [&](const __hip_bfloat16 *Ar, const double *xr) -> double {
    double a = 0.0;
    if (vec) {
      for (int j = lane * 8; j < n; j += 512) {
        float av[8];
        load8(Ar + j, av);
#pragma unroll
        for (int u = 0; u < 8; ++u) a += (double)av[u] * xr[j + u];
      }
    } else {
      for (int j = lane; j < n; j += 64) a += (double)a_val(Ar, j) * xr[j];
    }
    return a;
  };
  if (row0 / pm != (row0 + nr - 1) / pm || !vec) {
    for (int r = 0; r < nr; ++r) {
      const int row = row0 + r;
      const double t = wave_sum(row_sum(A + (long long)(row % pm) * n, x + gd.off_n + (long long)(row / pm) * n));
      if (lane == 0) y[gd.off_m + row] = t;
    }
    return;
  }
  const __hip_bfloat16 *Ab = A + (long long)(row0 % pm) * n;
  const double *xg = x + gd.off_n + (long long)(row0 / pm) * n;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = lane * 8; j < n; j += 512) {
    double xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = xg[j + u];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (r < nr) {
        float av[8];
        load8(Ab + (long long)r * n + j, av);
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[r] += (double)av[u] * xv[u];
      }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double t = wave_sum(acc[r]);
    if (lane == 0 && r < nr) y[gd.off_m + row0 + r] = t;
  }
}

// A^T x for a bf16 A and double vectors: the chunks of k_dense_bwd_part_d
// (the same rows per chunk, each column's rows in order), eight consecutive
// columns per thread, from one 16-byte load per row when the rows are vector
// aligned (a_vec: the 8 columns then lie in one layer), else column by
// column; k_dense_bwd_reduce_d adds the chunks.
__global__ __launch_bounds__(kBlock) void k_dense_bwd_part_db(
    const GroupDesc *G, const double *x, double *scratch, const InstScal *S,
    uint32_t mask, long long sstride) {
  const GroupDesc &gd = G[blockIdx.z];
  const int n = gd.plan.n, pm = gd.plan.m;
  const int c0 = (blockIdx.x * kBlock + threadIdx.x) * 8;
  if (c0 >= gd.n) return;
  if (mask && (S[blockIdx.z].flags & mask) != mask) return;
  const int nch = min(kDenseChunks, (pm + 63) / 64);
  const int c = blockIdx.y;
  if (c >= nch) return;
  const int R = (pm + nch - 1) / nch;
  const int i0 = c * R, i1 = min(pm, i0 + R);
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int w = min(8, gd.n - c0);
  double *out = scratch + blockIdx.z * sstride + (long long)c * gd.n + c0;
  if (!(gd.plan.a_vec && w == 8)) {
    for (int u = 0; u < w; ++u) {
      const int col = c0 + u;
      const __hip_bfloat16 *Ac = (const __hip_bfloat16 *)gd.plan.A + col % n;
      const double *xc = x + gd.off_m + (long long)(col / n) * pm;
      double a = 0.0;
      for (int i = i0; i < i1; ++i) a += (double)a_val(Ac, (long long)i * n) * xc[i];
      out[u] = a;
    }
    return;
  }
  const __hip_bfloat16 *A = (const __hip_bfloat16 *)gd.plan.A + c0 % n;
  const double *xg = x + gd.off_m + (long long)(c0 / n) * pm;
  int i = i0;
  for (; i + 4 <= i1; i += 4) {
    float a[4][8];
    double xi[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      load8(A + (long long)(i + r) * n, a[r]);
      xi[r] = xg[i + r];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] += (double)a[r][u] * xi[r];
  }
  for (; i < i1; ++i) {
    float a[8];
    load8(A + (long long)i * n, a);
    const double xi = xg[i];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += (double)a[u] * xi;
  }
#pragma unroll
  for (int u = 0; u < 8; ++u) out[u] = acc[u];
}

template <>
void launch_fwd<1>(hipStream_t s, const GroupDesc *G_dev, int G, int max_m,
                   int kind, int dtype, const double *x, double *y,
                   const InstScal *S, uint32_t mask) {
  if (kind == kDense && dtype == TV3_BF16) {
    hipLaunchKernelGGL(k_dense_fwd_db, dim3((max_m + 4 * (kBlock / 64) - 1) / (4 * (kBlock / 64)), G), dim3(kBlock), 0,
                       s, G_dev, x, y, S, mask);
    return;
  }
  if (kind == kDense) {
    hipLaunchKernelGGL(k_fwd_w1d<64>, dim3((max_m + kBlock / 64 - 1) / (kBlock / 64), G), dim3(kBlock), 0, s,
                       G_dev, x, y, S, mask);
    return;
  }
  hipLaunchKernelGGL(k_fwd_w1d<kSubF>, dim3((max_m + kBlock / kSubF - 1) / (kBlock / kSubF), G), dim3(kBlock), 0,
                     s, G_dev, x, y, S, mask);
}

template <>
void launch_bwd<1>(hipStream_t s, const GroupDesc *G_dev, int G, int max_n,
                   int kind, int dtype, const double *x, double *y,
                   const InstScal *S, uint32_t mask, const double *sub,
                   double *scratch, int submat_mb) {
  (void)submat_mb;
  if (kind == kDense) {
    // scratch holds kDenseChunks * max_n doubles per group
    const long long sstride = (long long)kDenseChunks * max_n;
    if (dtype == TV3_BF16)
      hipLaunchKernelGGL(k_dense_bwd_part_db, dim3((max_n + 8 * kBlock - 1) / (8 * kBlock), kDenseChunks, G),
                         dim3(kBlock), 0, s, G_dev, x, scratch, S, mask, sstride);
    else
      hipLaunchKernelGGL(k_dense_bwd_part_d, dim3((max_n + kBlock - 1) / kBlock, kDenseChunks, G), dim3(kBlock), 0,
                         s, G_dev, x, scratch, S, mask, sstride);
    hipLaunchKernelGGL(k_dense_bwd_reduce_d, dim3((max_n + kBlock - 1) / kBlock, G), dim3(kBlock), 0, s, G_dev,
                       (const double *)scratch, y, S, mask, sub, sstride);
    return;
  }
  hipLaunchKernelGGL(k_bwd_w1d<kSubT>, dim3((max_n + kBlock / kSubT - 1) / (kBlock / kSubT), G), dim3(kBlock), 0,
                     s, G_dev, x, y, S, mask, sub);
}

// ---------------------------------------------------------------------------
// Complex vectors, W = 1: the complex overloads (mat<complex<T>> U with
// mat<complex<T>>, sparse_mat<complex<T>> or geometry A, 2d/tval3_cpu_2d.cpp:
// 31-41, 58-68, 85-95).  y = A x takes the terms a x and y = A^H x
// (CblasConjTrans, 2d:308, 392, 774) the terms conj(a) x; each component of
// a term is one FP64 rounding of two FP64 products (exact for float
// components), summed in FP64 and rounded once.  The projector's weights are
// real: each component times the scale factor in the component type.
// ---------------------------------------------------------------------------
template <class C>
__device__ __forceinline__ void cterm(double &sr, double &si, C a, C v, bool conj) {
  if (!conj) {
    sr += (double)a.x * (double)v.x - (double)a.y * (double)v.y;
    si += (double)a.x * (double)v.y + (double)a.y * (double)v.x;
  } else {
    sr += (double)a.x * (double)v.x + (double)a.y * (double)v.y;
    si += (double)a.x * (double)v.y - (double)a.y * (double)v.x;
  }
}

template <int L, class C>
__global__ __launch_bounds__(kBlock) void k_fwd_w1c(const GroupDesc *G, const C *x, C *y,
                                                    const InstScal *S, uint32_t mask) {
  typedef decltype(C::x) R;
  const GroupDesc &gd = G[blockIdx.y];
  const int sl = threadIdx.x & (L - 1);
  const int row = blockIdx.x * (kBlock / L) + (threadIdx.x / L);
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const PlanDev &P = gd.plan;
  const int ar = row % P.m;
  const C *xg = x + gd.off_n + (long long)(row / P.m) * P.n;
  double sr = 0.0, si = 0.0;
  if (row < gd.m) {
    if (P.kind == kGeometry) {
      const RayDesc rd = P.rays[ar];
      const R sc = (R)P.scale;
      for (int j = sl; j <= rd.L; j += L) {
        const C v = xg[ray_lin(rd, j)];
        sr += (double)(v.x * sc);
        si += (double)(v.y * sc);
      }
    } else if (P.kind == kSparse) {
      const C *val = (const C *)P.r_valx;
      const int k1 = P.r_ptr[ar + 1];
      for (int k = P.r_ptr[ar] + sl; k < k1; k += L) cterm(sr, si, val[k], xg[P.r_ind[k]], false);
    } else {
      const C *A = (const C *)P.A + (long long)ar * P.n;
      for (int j = sl; j < P.n; j += L) cterm(sr, si, A[j], xg[j], false);
    }
  }
  sr = lane_sum<L>(sr);
  si = lane_sum<L>(si);
  if (sl == 0 && row < gd.m) y[gd.off_m + (long long)(row / P.m) * P.m + ar] = C((R)sr, (R)si);
}

template <int L, class C>
__global__ __launch_bounds__(kBlock) void k_bwd_w1c(const GroupDesc *G, const C *x, C *y,
                                                    const InstScal *S, uint32_t mask, const C *sub) {
  typedef decltype(C::x) R;
  const GroupDesc &gd = G[blockIdx.y];
  const int sl = threadIdx.x & (L - 1);
  const int col = blockIdx.x * (kBlock / L) + (threadIdx.x / L);
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const PlanDev &P = gd.plan;
  const bool geo = P.kind == kGeometry;
  const bool tiled = geo && P.t_col;
  int ac = col % P.n;
  const C *xg = x + gd.off_m + (long long)(col / P.n) * P.m;
  double sr = 0.0, si = 0.0;
  if (col < gd.n) {
    const int *ptr = tiled ? P.t_ptr : geo ? P.g_cptr : P.c_ptr;
    const int *ind = tiled ? P.t_ray : geo ? P.g_cray : P.c_ind;
    const int k0 = ptr[ac], k1 = ptr[ac + 1];
    if (tiled) ac = P.t_col[col];
    const R sc = (R)P.scale;
    const C *val = (const C *)P.c_valx;
    for (int k = k0 + sl; k < k1; k += L) {
      const C v = xg[ind[k]];
      if (geo) {
        sr += (double)(v.x * sc);
        si += (double)(v.y * sc);
      } else {
        cterm(sr, si, val[k], v, true);
      }
    }
  }
  sr = lane_sum<L>(sr);
  si = lane_sum<L>(si);
  if (sl == 0 && col < gd.n) {
    const long long o = gd.off_n + (long long)(col / P.n) * P.n + ac;
    const C v((R)sr, (R)si);
    y[o] = sub ? v - sub[o] : v;
  }
}

// dense A^H x: chunk c of a group covers rows [c*R, min(m, (c+1)*R)); a
// thread owns one column and writes its FP64 chunk sums (re, im)
template <class C>
__global__ __launch_bounds__(kBlock) void k_cdense_bwd_part(const GroupDesc *G, const C *x,
                                                            double *scratch, const InstScal *S,
                                                            uint32_t mask, long long sstride) {
  const GroupDesc &gd = G[blockIdx.z];
  const int n = gd.plan.n, pm = gd.plan.m;
  const int col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= gd.n) return;
  if (mask && (S[blockIdx.z].flags & mask) != mask) return;
  const int nch = min(kDenseChunks, (pm + 63) / 64);
  const int c = blockIdx.y;
  if (c >= nch) return;
  const int Rr = (pm + nch - 1) / nch;
  const int i0 = c * Rr, i1 = min(pm, i0 + Rr);
  const C *A = (const C *)gd.plan.A + col % n;
  const C *xg = x + gd.off_m + (long long)(col / n) * pm;
  double sr = 0.0, si = 0.0;
  for (int i = i0; i < i1; ++i) cterm(sr, si, A[(long long)i * n], xg[i], true);
  double *out = scratch + blockIdx.z * sstride + 2 * ((long long)c * gd.n + col);
  out[0] = sr;
  out[1] = si;
}

template <class C>
__global__ __launch_bounds__(kBlock) void k_cdense_bwd_reduce(const GroupDesc *G, const double *scratch,
                                                              C *y, const InstScal *S, uint32_t mask,
                                                              const C *sub, long long sstride) {
  typedef decltype(C::x) R;
  const GroupDesc &gd = G[blockIdx.y];
  const int col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= gd.n) return;
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const int nch = min(kDenseChunks, (gd.plan.m + 63) / 64);
  const double *sp = scratch + blockIdx.y * sstride + 2 * (long long)col;
  double sr = 0.0, si = 0.0;
  for (int c = 0; c < nch; ++c) {
    sr += sp[2 * (long long)c * gd.n];
    si += sp[2 * (long long)c * gd.n + 1];
  }
  const long long o = gd.off_n + col;
  const C v((R)sr, (R)si);
  y[o] = sub ? v - sub[o] : v;
}

// get_g's blocked A^H (A U) (2d/tval3_cpu_2d.cpp:361-394) with complex
// entries, one layer: a thread owns one column and walks the rows in order;
// FP64 within a block of rows, each block sum rounded per component and
// added in the component type in block order.  y = g - sub.
template <class C>
__global__ __launch_bounds__(kBlock) void k_cdense_bwd_blk(const GroupDesc *G, const C *x, C *y,
                                                           const InstScal *S, uint32_t mask,
                                                           const C *sub, int submat_mb) {
  typedef decltype(C::x) R;
  const GroupDesc &gd = G[blockIdx.y];
  const int col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= gd.n) return;
  if (mask && (S[blockIdx.y].flags & mask) != mask) return;
  const int n = gd.plan.n, m = gd.plan.m;
  const int rb = dense_block_rows(submat_mb, n, (int)sizeof(C));
  const C *A = (const C *)gd.plan.A + col;
  const C *xg = x + gd.off_m;
  R gr = 0, gi = 0;
  for (int y0 = 0; y0 < m; y0 += rb) {
    const int y1 = min(m, y0 + rb);
    double sr = 0.0, si = 0.0;
    for (int i = y0; i < y1; ++i) cterm(sr, si, A[(long long)i * n], xg[i], true);
    gr = gr + (R)sr;
    gi = gi + (R)si;
  }
  const long long o = gd.off_n + col;
  const C v(gr, gi);
  y[o] = sub ? v - sub[o] : v;
}

template <class C>
static void launch_fwd_c(hipStream_t s, const GroupDesc *G_dev, int G, int max_m, int kind,
                         const C *x, C *y, const InstScal *S, uint32_t mask) {
  if (kind == kDense) {
    hipLaunchKernelGGL((k_fwd_w1c<64, C>), dim3((max_m + kBlock / 64 - 1) / (kBlock / 64), G), dim3(kBlock), 0,
                       s, G_dev, x, y, S, mask);
    return;
  }
  hipLaunchKernelGGL((k_fwd_w1c<kSubF, C>), dim3((max_m + kBlock / kSubF - 1) / (kBlock / kSubF), G),
                     dim3(kBlock), 0, s, G_dev, x, y, S, mask);
}

template <class C>
static void launch_bwd_c(hipStream_t s, const GroupDesc *G_dev, int G, int max_n, int kind,
                         const C *x, C *y, const InstScal *S, uint32_t mask, const C *sub,
                         C *scratch, int submat_mb) {
  if (kind == kDense && submat_mb > 0) {
    hipLaunchKernelGGL((k_cdense_bwd_blk<C>), dim3((max_n + kBlock - 1) / kBlock, G), dim3(kBlock), 0, s,
                       G_dev, x, y, S, mask, sub, submat_mb);
    return;
  }
  if (kind == kDense) {
    // scratch holds 2 * kDenseChunks * max_n doubles per group
    const long long sstride = 2LL * kDenseChunks * max_n;
    hipLaunchKernelGGL((k_cdense_bwd_part<C>), dim3((max_n + kBlock - 1) / kBlock, kDenseChunks, G),
                       dim3(kBlock), 0, s, G_dev, x, (double *)scratch, S, mask, sstride);
    hipLaunchKernelGGL((k_cdense_bwd_reduce<C>), dim3((max_n + kBlock - 1) / kBlock, G), dim3(kBlock), 0, s,
                       G_dev, (const double *)scratch, y, S, mask, sub, sstride);
    return;
  }
  hipLaunchKernelGGL((k_bwd_w1c<kSubT, C>), dim3((max_n + kBlock / kSubT - 1) / (kBlock / kSubT), G),
                     dim3(kBlock), 0, s, G_dev, x, y, S, mask, sub);
}

template <>
void launch_fwd<1>(hipStream_t s, const GroupDesc *G_dev, int G, int max_m, int kind, int dtype,
                   const cfloat *x, cfloat *y, const InstScal *S, uint32_t mask) {
  (void)dtype;
  launch_fwd_c(s, G_dev, G, max_m, kind, x, y, S, mask);
}
template <>
void launch_fwd<1>(hipStream_t s, const GroupDesc *G_dev, int G, int max_m, int kind, int dtype,
                   const cdouble *x, cdouble *y, const InstScal *S, uint32_t mask) {
  (void)dtype;
  launch_fwd_c(s, G_dev, G, max_m, kind, x, y, S, mask);
}
template <>
void launch_bwd<1>(hipStream_t s, const GroupDesc *G_dev, int G, int max_n, int kind, int dtype,
                   const cfloat *x, cfloat *y, const InstScal *S, uint32_t mask, const cfloat *sub,
                   cfloat *scratch, int submat_mb) {
  (void)dtype;
  launch_bwd_c(s, G_dev, G, max_n, kind, x, y, S, mask, sub, scratch, submat_mb);
}
template <>
void launch_bwd<1>(hipStream_t s, const GroupDesc *G_dev, int G, int max_n, int kind, int dtype,
                   const cdouble *x, cdouble *y, const InstScal *S, uint32_t mask, const cdouble *sub,
                   cdouble *scratch, int submat_mb) {
  (void)dtype;
  launch_bwd_c(s, G_dev, G, max_n, kind, x, y, S, mask, sub, scratch, submat_mb);
}

// ---------------------------------------------------------------------------
// geometry plan construction
// ---------------------------------------------------------------------------

// Writes the voxel list of every row (walk order) at lin[ptr[row] ...].
__global__ void k_walk_export(const RayDesc *rays, int m, const long long *ptr,
                              int *lin) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (row >= m) return;
  const RayDesc rd = rays[row];
  const long long o = ptr[row];
  for (int j = lane; j <= rd.L; j += 64) lin[o + j] = ray_lin(rd, j);
}

}  // namespace tv3

// ===========================================================================
// C-ABI: context, options, plans
// ===========================================================================
extern "C" {

void tv3_options_default(tv3_options *o) {
  // Public TVAL3 defaults (the reference's values sit in tval3_types.h,
  // which its tree does not ship).
  o->mu = 256.0; o->beta = 32.0; o->mu0 = 256.0; o->beta0 = 32.0;
  o->tol = 1e-4; o->tol_inn = 1e-3; o->maxit = 300; o->maxcnt = 10;
  o->gam = 0.9995; o->gamma = 0.6; o->rate_gam = 0.9; o->rate_cnt = 2.0;
  o->c = 1e-5; o->nonneg = 0; o->isreal = 1; o->numInitIterations = 0;
  o->skipMulRatio = 0.0; o->maxRelativeChange = 0.0;
  o->scal_double = 0; o->tv_iso = 0; o->dt_exact = 0; o->ls_exact = 0;
  o->graph = 0;
  o->submat_mb = 5;
}

int tv3_ctx_create(int device, void *stream, tv3_ctx **out) {
  if (!out) return TV3_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TV3_ERR_NOGPU;
  if (device < 0 || device >= ndev) return TV3_ERR_ARG;
  tv3_ctx *c = new tv3_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete c; return TV3_ERR_HIP; }
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return TV3_ERR_HIP;
    }
    c->own_stream = true;
  }
  // the no-allocation gemv's device scratch of this GPU (once per device)
  if (tv3::noalloc_attach(device) != TV3_OK) {
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TV3_ERR_HIP;
  }
  if (hipHostMalloc(&c->pinned, 1 << 16, hipHostMallocDefault) != hipSuccess) {
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TV3_ERR_ALLOC;
  }
  c->pinned_bytes = 1 << 16;
  (void)hipEventCreate(&c->ev0);
  (void)hipEventCreate(&c->ev1);
  *out = c;
  return TV3_OK;
}

void tv3_ctx_destroy(tv3_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->ws) (void)hipFree(c->ws);
  if (c->prof_bytes) (void)hipFree(c->prof_bytes);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->hflags) (void)hipHostFree(c->hflags);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *tv3_ctx_error(const tv3_ctx *c) { return c ? c->err.c_str() : ""; }

int tv3_ctx_sync(tv3_ctx *c) {
  if (!c) return TV3_ERR_ARG;
  return hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

void *tv3_ctx_stream(const tv3_ctx *c) { return c ? (void *)c->stream : nullptr; }

int tv3_mem_alloc(tv3_ctx *c, size_t bytes, void **out) {
  if (!c || !out) return TV3_ERR_ARG;
  *out = nullptr;
  if (!bytes) return TV3_OK;
  int rc = hip_check(c, hipSetDevice(c->device), "hipSetDevice");
  if (rc) return rc;
  if (hipMalloc(out, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return set_err(c, TV3_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return TV3_OK;
}

int tv3_mem_free(tv3_ctx *c, void *p) {
  if (!c) return TV3_ERR_ARG;
  if (!p) return TV3_OK;
  int rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (rc) return rc;
  return hip_check(c, hipFree(p), "hipFree");
}

int tv3_mem_upload(tv3_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c || (bytes && (!dst || !src))) return TV3_ERR_ARG;
  if (!bytes) return TV3_OK;
  int rc = hip_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream), "upload");
  if (rc) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "upload sync");
}

int tv3_mem_download(tv3_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c || (bytes && (!dst || !src))) return TV3_ERR_ARG;
  if (!bytes) return TV3_OK;
  int rc = hip_check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "download");
  if (rc) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "download sync");
}

int tv3_ctx_profile(tv3_ctx *c, int enable) {
  if (!c) return TV3_ERR_ARG;
  (void)hipStreamSynchronize(c->stream);
  c->evkind.clear();
  c->prof = enable != 0;
  c->prof_level = enable;
  c->prof_open = false;
  for (int k = 0; k < kProfSlots; ++k) {
    c->prof_ms[k] = 0;
    c->prof_n[k] = 0;
  }
  for (unsigned long long &v : c->prof_hbytes) v = 0;
  if (c->prof) {
    if (!c->prof_bytes) {
      int rc = hip_check(c, hipMalloc(&c->prof_bytes, 4 * sizeof(unsigned long long)), "profile hipMalloc");
      if (rc) { c->prof = false; c->prof_bytes = nullptr; return rc; }
    }
    int rc = hip_check(c, hipMemsetAsync(c->prof_bytes, 0, 4 * sizeof(unsigned long long), c->stream), "profile clear");
    if (rc) { c->prof = false; return rc; }
  }
  return TV3_OK;
}

int tv3_ctx_profile_read(tv3_ctx *c, double *ms, long long *count, double *bytes) {
  if (!c || !ms || !count) return TV3_ERR_ARG;
  int rc = hip_check(c, hipStreamSynchronize(c->stream), "profile sync");
  if (rc) return rc;
  prof_collect(c);
  const int slot[2] = {kProfGemvN, kProfGemvT};
  for (int k = 0; k < 2; ++k) { ms[k] = c->prof_ms[slot[k]]; count[k] = c->prof_n[slot[k]]; }
  if (bytes) {
    unsigned long long hb[4] = {0, 0, 0, 0};
    if (c->prof_bytes) {
      rc = hip_check(c, hipMemcpy(hb, c->prof_bytes, sizeof(hb), hipMemcpyDeviceToHost), "profile read");
      if (rc) return rc;
    }
    bytes[0] = (double)(hb[0] + c->prof_hbytes[0]);
    bytes[1] = (double)(hb[1] + c->prof_hbytes[1]);
  }
  return TV3_OK;
}

int tv3_ctx_profile_lanes(tv3_ctx *c, long long *lanes) {
  if (!c || !lanes) return TV3_ERR_ARG;
  int rc = hip_check(c, hipStreamSynchronize(c->stream), "profile sync");
  if (rc) return rc;
  unsigned long long hb[4] = {0, 0, 0, 0};
  if (c->prof_bytes) {
    rc = hip_check(c, hipMemcpy(hb, c->prof_bytes, sizeof(hb), hipMemcpyDeviceToHost), "profile read");
    if (rc) return rc;
  }
  lanes[0] = (long long)(hb[2] + c->prof_hbytes[2]);
  lanes[1] = (long long)(hb[3] + c->prof_hbytes[3]);
  return TV3_OK;
}

int tv3_ctx_profile_slots(tv3_ctx *c, tv3_profile_slot *out, int cap, int *n) {
  if (!c || !n || (cap > 0 && !out)) return TV3_ERR_ARG;
  int rc = hip_check(c, hipStreamSynchronize(c->stream), "profile sync");
  if (rc) return rc;
  prof_collect(c);
  int k = 0;
  for (int sl = 0; sl < kProfSlots; ++sl) {
    if (c->prof_n[sl] == 0) continue;
    if (k < cap) {
      tv3_profile_slot &o = out[k];
      o.slot = sl;
      o.runs = c->prof_n[sl];
      o.ms = c->prof_ms[sl];
      std::snprintf(o.name, sizeof(o.name), "%s", prof_slot_name(sl));
    }
    ++k;
  }
  *n = k;
  return TV3_OK;
}

int tv3_plan_dense(tv3_ctx *ctx, int m, int n, int dtype, const void *A,
                   tv3_plan **out) {
  if (!ctx || !out) return TV3_ERR_ARG;
  if (m <= 0 || n <= 0 || !A || dtype < TV3_F32 || dtype > TV3_C128)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: dense plan needs m, n > 0, A and dtype f32/bf16/f64/c64/c128");
  if ((dtype == TV3_C64 && (uintptr_t)A % 8) || (dtype == TV3_C128 && (uintptr_t)A % 16))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: complex A must be aligned to its element size");
  tv3_plan *p = new tv3_plan();
  p->ctx = ctx;
  p->dev.kind = kDense; p->dev.m = m; p->dev.n = n; p->dev.A = A;
  p->dev.dtype = dtype;
  p->dev.a_vec = ((uintptr_t)A % 16 == 0) && (n % 8 == 0);
  p->nnz = (long long)m * n;
  p->dev.nnz = p->nnz;
  *out = p;
  return TV3_OK;
}

int tv3_plan_sparse(tv3_ctx *ctx, int m, int n, int nnz, const int *csr_ptr,
                    const int *csr_ind, const float *csr_val,
                    const int *csc_ptr, const int *csc_ind,
                    const float *csc_val, tv3_plan **out) {
  if (!ctx || !out) return TV3_ERR_ARG;
  if (m <= 0 || n <= 0 || nnz < 0)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: sparse plan needs m, n > 0");
  if (!csr_ptr || !csc_ptr || (nnz > 0 && (!csr_ind || !csr_val || !csc_ind || !csc_val)))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: A must be available in CSR and CSC format!");
  tv3_plan *p = new tv3_plan();
  p->ctx = ctx;
  PlanDev &d = p->dev;
  d.kind = kSparse; d.m = m; d.n = n;
  d.r_ptr = csr_ptr; d.r_ind = csr_ind; d.r_val = csr_val;
  d.c_ptr = csc_ptr; d.c_ind = csc_ind; d.c_val = csc_val;
  p->nnz = nnz;
  p->dev.nnz = nnz;
  *out = p;
  return TV3_OK;
}

int tv3_plan_sparse_f64(tv3_ctx *ctx, int m, int n, int nnz, const int *csr_ptr,
                        const int *csr_ind, const double *csr_val,
                        const int *csc_ptr, const int *csc_ind,
                        const double *csc_val, tv3_plan **out) {
  if (!ctx || !out) return TV3_ERR_ARG;
  if (m <= 0 || n <= 0 || nnz < 0)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: sparse plan needs m, n > 0");
  if (!csr_ptr || !csc_ptr || (nnz > 0 && (!csr_ind || !csr_val || !csc_ind || !csc_val)))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: A must be available in CSR and CSC format!");
  tv3_plan *p = new tv3_plan();
  p->ctx = ctx;
  PlanDev &d = p->dev;
  d.kind = kSparse; d.m = m; d.n = n;
  d.r_ptr = csr_ptr; d.r_ind = csr_ind; d.r_vald = csr_val;
  d.c_ptr = csc_ptr; d.c_ind = csc_ind; d.c_vald = csc_val;
  p->nnz = nnz;
  p->dev.nnz = nnz;
  *out = p;
  return TV3_OK;
}

static int plan_sparse_c(tv3_ctx *ctx, int m, int n, int nnz, const int *csr_ptr,
                         const int *csr_ind, const void *csr_val, const int *csc_ptr,
                         const int *csc_ind, const void *csc_val, int cval, tv3_plan **out) {
  if (!ctx || !out) return TV3_ERR_ARG;
  if (m <= 0 || n <= 0 || nnz < 0)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: sparse plan needs m, n > 0");
  if (!csr_ptr || !csc_ptr || (nnz > 0 && (!csr_ind || !csr_val || !csc_ind || !csc_val)))
    return set_err(ctx, TV3_ERR_ARG, "Argument error: A must be available in CSR and CSC format!");
  const uintptr_t al = cval == 2 ? 8 : 16;
  if ((uintptr_t)csr_val % al || (uintptr_t)csc_val % al)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: complex values must be aligned to their element size");
  tv3_plan *p = new tv3_plan();
  p->ctx = ctx;
  PlanDev &d = p->dev;
  d.kind = kSparse; d.m = m; d.n = n;
  d.r_ptr = csr_ptr; d.r_ind = csr_ind; d.r_valx = csr_val;
  d.c_ptr = csc_ptr; d.c_ind = csc_ind; d.c_valx = csc_val;
  d.cval = cval;
  p->nnz = nnz;
  p->dev.nnz = nnz;
  *out = p;
  return TV3_OK;
}

int tv3_plan_sparse_c64(tv3_ctx *ctx, int m, int n, int nnz, const int *csr_ptr,
                        const int *csr_ind, const float *csr_val, const int *csc_ptr,
                        const int *csc_ind, const float *csc_val, tv3_plan **out) {
  return plan_sparse_c(ctx, m, n, nnz, csr_ptr, csr_ind, csr_val, csc_ptr, csc_ind, csc_val, 2, out);
}

int tv3_plan_sparse_c128(tv3_ctx *ctx, int m, int n, int nnz, const int *csr_ptr,
                         const int *csr_ind, const double *csr_val, const int *csc_ptr,
                         const int *csc_ind, const double *csc_val, tv3_plan **out) {
  return plan_sparse_c(ctx, m, n, nnz, csr_ptr, csr_ind, csr_val, csc_ptr, csc_ind, csc_val, 3, out);
}

int tv3_plan_geometry(tv3_ctx *ctx, const tv3_geometry *g, tv3_plan **out) {
  if (!ctx || !g || !out) return TV3_ERR_ARG;
  *out = nullptr;
  const int E = g->num_emitters, R = g->num_receivers;
  if (g->num_mps <= 0 || E <= 0 || R <= 0 || g->num_paths <= 0 ||
      g->rv_x <= 0 || g->rv_y <= 0 || g->rv_z <= 0 || !g->use_path)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: invalid geometry");
  const long long nvox = (long long)g->rv_x * g->rv_y * g->rv_z;
  const long long m = (long long)g->num_mps * g->num_paths;
  if (nvox > 0x7fffffffLL || m > 0x7fffffffLL)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: geometry too large");
  std::vector<RayDesc> rays(m);
  for (auto &r : rays) { r = RayDesc{}; r.L = -1; }
  const int sx = 1, sy = g->rv_x, sz = g->rv_x * g->rv_y;
  auto inside = [&](int x, int y, int z) {
    return x >= 0 && x < g->rv_x && y >= 0 && y < g->rv_y && z >= 0 && z < g->rv_z;
  };
  for (int i = 0; i < g->num_mps; ++i)
    for (int e = 0; e < E; ++e)
      for (int r = 0; r < R; ++r) {
        const int up = g->use_path[e * R + r];
        if (up == 0) continue;
        if (up < 0 || up > g->num_paths)
          return set_err(ctx, TV3_ERR_ARG, "Argument error: use_path entry out of range");
        const long long row = (long long)i * g->num_paths + up - 1;
        if (rays[row].L >= 0)
          return set_err(ctx, TV3_ERR_ARG, "Argument error: two pairs map to one measurement row");
        const int x0 = g->x_emitters[i * E + e], y0 = g->y_emitters[i * E + e],
                  z0 = g->z_emitters[i * E + e];
        const int x1 = g->x_receivers[i * R + r], y1 = g->y_receivers[i * R + r],
                  z1 = g->z_receivers[i * R + r];
        if (!inside(x0, y0, z0) || !inside(x1, y1, z1))
          return set_err(ctx, TV3_ERR_ARG, "Argument error: transducer outside the volume");
        const int dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
        const int mx = std::abs(dx), my = std::abs(dy), mz = std::abs(dz);
        const int ix = dx < 0 ? -sx : sx, iy = dy < 0 ? -sy : sy, iz = dz < 0 ? -sz : sz;
        RayDesc d;
        d.lin0 = x0 * sx + y0 * sy + z0 * sz;
        // major axis: the first of x, y, z with the largest |d|
        if (mx >= my && mx >= mz) {
          d.L = mx; d.s_major = ix; d.s_a = iy; d.s_b = iz; d.ma = my; d.mb = mz;
        } else if (my >= mx && my >= mz) {
          d.L = my; d.s_major = iy; d.s_a = ix; d.s_b = iz; d.ma = mx; d.mb = mz;
        } else {
          d.L = mz; d.s_major = iz; d.s_a = ix; d.s_b = iy; d.ma = mx; d.mb = my;
        }
        // the closed-form quotient is exact while 2*m*j + L < 2^24
        if (d.L > 2048)
          return set_err(ctx, TV3_ERR_ARG, "Argument error: ray longer than 2048 voxels");
        d.inv2L = d.L > 0 ? 1.0f / (float)(2 * d.L) : 0.0f;
        rays[row] = d;
      }
  // forward walk order: tiles of te emitters x tr receivers per measurement
  // plane, each tile in pair order
  std::vector<int> order;
  {
    int te, tr;
    ray_tile(te, tr);
    if (te > 0 && tr > 0) {
      order.reserve(m);
      for (int i = 0; i < g->num_mps; ++i)
        for (int e0 = 0; e0 < E; e0 += te)
          for (int r0 = 0; r0 < R; r0 += tr)
            for (int e = e0; e < std::min(E, e0 + te); ++e)
              for (int r = r0; r < std::min(R, r0 + tr); ++r) {
                const int up = g->use_path[e * R + r];
                if (up != 0) order.push_back((int)((long long)i * g->num_paths + up - 1));
              }
      // rows no pair maps to (L = -1) keep their place at the end
      std::vector<char> seen(m, 0);
      for (int r : order) seen[r] = 1;
      for (long long r = 0; r < m; ++r)
        if (!seen[r]) order.push_back((int)r);
    }
  }
  // transpose walk order: t x t tiles of each slice (raster order for t = 0)
  std::vector<int> corder;
  corder.reserve(nvox);
  if (g->rv_z > 1 && vox_brick() > 0) {
    const int b = vox_brick();
    for (int z0 = 0; z0 < g->rv_z; z0 += b)
      for (int y0 = 0; y0 < g->rv_y; y0 += b)
        for (int x0 = 0; x0 < g->rv_x; x0 += b)
          for (int z = z0; z < std::min(g->rv_z, z0 + b); ++z)
            for (int y = y0; y < std::min(g->rv_y, y0 + b); ++y)
              for (int x = x0; x < std::min(g->rv_x, x0 + b); ++x)
                corder.push_back(x * sx + y * sy + z * sz);
  } else {
    const int t = vox_tile() > 0 ? vox_tile() : std::max(g->rv_x, g->rv_y);
    for (int z = 0; z < g->rv_z; ++z)
      for (int y0 = 0; y0 < g->rv_y; y0 += t)
        for (int x0 = 0; x0 < g->rv_x; x0 += t)
          for (int y = y0; y < std::min(g->rv_y, y0 + t); ++y)
            for (int x = x0; x < std::min(g->rv_x, x0 + t); ++x)
              corder.push_back(x * sx + y * sy + z * sz);
  }
  tv3_plan *p = new tv3_plan();
  p->ctx = ctx;
  PlanDev &dv = p->dev;
  dv.kind = kGeometry; dv.m = (int)m; dv.n = (int)nvox; dv.scale = g->scale_factor;
  p->walk_ptr.assign(m + 1, 0);
  for (long long r = 0; r < m; ++r)
    p->walk_ptr[r + 1] = p->walk_ptr[r] + (rays[r].L >= 0 ? rays[r].L + 1 : 0);
  const long long nnz = p->walk_ptr[m];
  p->nnz = nnz;
  p->dev.nnz = nnz;
  if (nnz > 0x7fffffffLL) { delete p; return set_err(ctx, TV3_ERR_ARG, "Argument error: too many ray voxels"); }
  hipStream_t s = ctx->stream;
  int rc = TV3_OK;
  long long *d_ptr = nullptr;
  int *d_lin = nullptr;
  std::vector<int> lin(nnz);
#define TV3_TRY(x, what) do { rc = hip_check(ctx, (x), what); if (rc) goto fail; } while (0)
  TV3_TRY(hipMalloc(&p->rays, sizeof(RayDesc) * m), "hipMalloc rays");
  TV3_TRY(hipMemcpyAsync(p->rays, rays.data(), sizeof(RayDesc) * m, hipMemcpyHostToDevice, s), "copy rays");
  if (!order.empty()) {
    TV3_TRY(hipMalloc(&p->order, sizeof(int) * m), "hipMalloc ray order");
    TV3_TRY(hipMemcpyAsync(p->order, order.data(), sizeof(int) * m, hipMemcpyHostToDevice, s), "copy ray order");
  }

  TV3_TRY(hipMalloc(&d_ptr, sizeof(long long) * (m + 1)), "hipMalloc walk ptr");
  TV3_TRY(hipMalloc(&d_lin, sizeof(int) * std::max<long long>(nnz, 1)), "hipMalloc walk");
  TV3_TRY(hipMemcpyAsync(d_ptr, p->walk_ptr.data(), sizeof(long long) * (m + 1), hipMemcpyHostToDevice, s), "copy walk ptr");
  hipLaunchKernelGGL(k_walk_export, dim3((unsigned)((m + 3) / 4)), dim3(kBlock), 0, s, p->rays, (int)m, d_ptr, d_lin);
  TV3_TRY(hipGetLastError(), "k_walk_export");
  if (nnz > 0)
    TV3_TRY(hipMemcpyAsync(lin.data(), d_lin, sizeof(int) * nnz, hipMemcpyDeviceToHost, s), "copy walk");
  TV3_TRY(hipStreamSynchronize(s), "walk sync");
  {
    // voxel-major transpose, rays ascending within each voxel (counting sort)
    std::vector<int> cptr(nvox + 1, 0), cray(std::max<long long>(nnz, 1));
    for (long long k = 0; k < nnz; ++k) cptr[lin[k] + 1]++;
    for (long long v = 0; v < nvox; ++v) cptr[v + 1] += cptr[v];
    std::vector<int> fill(cptr.begin(), cptr.end() - 1);
    for (long long r = 0; r < m; ++r)
      for (long long k = p->walk_ptr[r]; k < p->walk_ptr[r + 1]; ++k)
        cray[fill[lin[k]]++] = (int)r;
    TV3_TRY(hipMalloc(&p->cptr, sizeof(int) * (nvox + 1)), "hipMalloc cptr");
    TV3_TRY(hipMalloc(&p->cray, sizeof(int) * cray.size()), "hipMalloc cray");
    TV3_TRY(hipMemcpyAsync(p->cptr, cptr.data(), sizeof(int) * (nvox + 1), hipMemcpyHostToDevice, s), "copy cptr");
    TV3_TRY(hipMemcpyAsync(p->cray, cray.data(), sizeof(int) * cray.size(), hipMemcpyHostToDevice, s), "copy cray");
    // voxels no ray crosses go to the end of the transpose order (stable);
    // 3-D volumes only: on the 256 ring (21% uncrossed) it left both
    // transposes' times unchanged and raised the 64-lane one's HBM fetches
    // by 15%
    int t_nz = (int)nvox;
    if (g->rv_z > 1 && vox_nz_first()) {
      auto mid = std::stable_partition(corder.begin(), corder.end(),
                                       [&](int v) { return cptr[v + 1] > cptr[v]; });
      t_nz = (int)(mid - corder.begin());
    }
    dv.t_nz = t_nz;
    // the same lists in the walk order of the 64-lane transpose
    std::vector<int> tptr(nvox + 1, 0), tray(cray.size());
    for (long long k = 0; k < nvox; ++k) {
      const int v = corder[k];
      tptr[k + 1] = tptr[k] + (cptr[v + 1] - cptr[v]);
      std::copy(cray.begin() + cptr[v], cray.begin() + cptr[v + 1], tray.begin() + tptr[k]);
    }
    TV3_TRY(hipMalloc(&p->tptr, sizeof(int) * (nvox + 1)), "hipMalloc tptr");
    TV3_TRY(hipMalloc(&p->tray, sizeof(int) * tray.size()), "hipMalloc tray");
    TV3_TRY(hipMalloc(&p->tcol, sizeof(int) * nvox), "hipMalloc tcol");
    TV3_TRY(hipMemcpyAsync(p->tptr, tptr.data(), sizeof(int) * (nvox + 1), hipMemcpyHostToDevice, s), "copy tptr");
    TV3_TRY(hipMemcpyAsync(p->tray, tray.data(), sizeof(int) * tray.size(), hipMemcpyHostToDevice, s), "copy tray");
    TV3_TRY(hipMemcpyAsync(p->tcol, corder.data(), sizeof(int) * nvox, hipMemcpyHostToDevice, s), "copy tcol");
    TV3_TRY(hipStreamSynchronize(s), "plan sync");
    // staged transpose (k_bwd_w1s): the crossed voxels brick by brick (8^3
    // bricks of a volume, 16^2 squares of a slice), then the uncrossed ones;
    // tiles of V slots, each with its ascending ray list, and the slots'
    // lists in slot order, each entry the ray's position in its tile's list
    {
      const bool vol = g->rv_z > 1;
      const int b = vol ? 8 : 16, bz = vol ? 8 : 1;
      const int V = kStageV;
      std::vector<int> scol;
      scol.reserve(nvox);
      for (int z0 = 0; z0 < g->rv_z; z0 += bz)
        for (int y0 = 0; y0 < g->rv_y; y0 += b)
          for (int x0 = 0; x0 < g->rv_x; x0 += b)
            for (int z = z0; z < std::min(g->rv_z, z0 + bz); ++z)
              for (int y = y0; y < std::min(g->rv_y, y0 + b); ++y)
                for (int x = x0; x < std::min(g->rv_x, x0 + b); ++x) {
                  const int v = x * sx + y * sy + z * sz;
                  const int len = cptr[v + 1] - cptr[v];
                  if (len > 0 && len <= kStageH) scol.push_back(v);
                }
      const int s_nz = (int)scol.size();
      // the voxels with longer lists, longest first, then the uncrossed ones
      std::vector<int> heavy;
      long long hnnz = 0;
      for (long long v = 0; v < nvox; ++v)
        if (cptr[v + 1] - cptr[v] > kStageH) {
          heavy.push_back((int)v);
          hnnz += cptr[v + 1] - cptr[v];
        }
      std::stable_sort(heavy.begin(), heavy.end(),
                       [&](int a, int c) { return cptr[a + 1] - cptr[a] > cptr[c + 1] - cptr[c]; });
      scol.insert(scol.end(), heavy.begin(), heavy.end());
      for (long long v = 0; v < nvox; ++v)
        if (cptr[v + 1] == cptr[v]) scol.push_back((int)v);
      const int T = (s_nz + V - 1) / V;
      std::vector<int> sptr(s_nz + 1, 0), srptr(T + 1, 0), sray, tl, pos(m, 0), seen(m, -1);
      std::vector<unsigned short> sloc;
      sloc.reserve(std::max<long long>(nnz, 1));
      long long dnnz = 0;
      for (int t = 0; t < T; ++t) {
        tl.clear();
        const int s1 = std::min(s_nz, (t + 1) * V);
        for (int k = t * V; k < s1; ++k) {
          const int v = scol[k];
          for (int e = cptr[v]; e < cptr[v + 1]; ++e)
            if (seen[cray[e]] != t) {
              seen[cray[e]] = t;
              tl.push_back(cray[e]);
            }
        }
        const bool st = (int)tl.size() <= stage_rays();
        if (st) {
          std::sort(tl.begin(), tl.end());
          for (int i = 0; i < (int)tl.size(); ++i) pos[tl[i]] = i;
          sray.insert(sray.end(), tl.begin(), tl.end());
        }
        srptr[t + 1] = (int)sray.size();
        for (int k = t * V; k < s1; ++k) {
          const int v = scol[k];
          for (int e = cptr[v]; e < cptr[v + 1]; ++e)
            sloc.push_back(st ? (unsigned short)pos[cray[e]] : (unsigned short)0);
          if (!st) dnnz += cptr[v + 1] - cptr[v];
          sptr[k + 1] = (int)sloc.size();
        }
      }
      if (sray.empty()) sray.push_back(0);
      if (sloc.empty()) sloc.push_back(0);
      dv.s_nz = s_nz;
      dv.s_nh = (int)heavy.size();
      dv.s_V = V;
      dv.s_rays = srptr[T];
      dv.s_dnnz = dnnz + hnnz;
      TV3_TRY(hipMalloc(&p->scol, sizeof(int) * nvox), "hipMalloc scol");
      TV3_TRY(hipMalloc(&p->sptr, sizeof(int) * (s_nz + 1)), "hipMalloc sptr");
      TV3_TRY(hipMalloc(&p->srptr, sizeof(int) * (T + 1)), "hipMalloc srptr");
      TV3_TRY(hipMalloc(&p->sray, sizeof(int) * sray.size()), "hipMalloc sray");
      TV3_TRY(hipMalloc(&p->sloc, sizeof(unsigned short) * sloc.size()), "hipMalloc sloc");
      TV3_TRY(hipMemcpyAsync(p->scol, scol.data(), sizeof(int) * nvox, hipMemcpyHostToDevice, s), "copy scol");
      TV3_TRY(hipMemcpyAsync(p->sptr, sptr.data(), sizeof(int) * (s_nz + 1), hipMemcpyHostToDevice, s), "copy sptr");
      TV3_TRY(hipMemcpyAsync(p->srptr, srptr.data(), sizeof(int) * (T + 1), hipMemcpyHostToDevice, s), "copy srptr");
      TV3_TRY(hipMemcpyAsync(p->sray, sray.data(), sizeof(int) * sray.size(), hipMemcpyHostToDevice, s), "copy sray");
      TV3_TRY(hipMemcpyAsync(p->sloc, sloc.data(), sizeof(unsigned short) * sloc.size(), hipMemcpyHostToDevice, s),
              "copy sloc");
      TV3_TRY(hipStreamSynchronize(s), "plan sync");
    }
  }
#undef TV3_TRY
  (void)hipFree(d_ptr);
  (void)hipFree(d_lin);
  dv.rays = p->rays; dv.g_cptr = p->cptr; dv.g_cray = p->cray; dv.ray_order = p->order;
  dv.t_ptr = p->tptr; dv.t_ray = p->tray; dv.t_col = p->tcol;
  dv.s_col = p->scol; dv.s_ptr = p->sptr; dv.s_rptr = p->srptr; dv.s_ray = p->sray; dv.s_loc = p->sloc;
  *out = p;
  return TV3_OK;
fail:
  if (d_ptr) (void)hipFree(d_ptr);
  if (d_lin) (void)hipFree(d_lin);
  if (p->rays) (void)hipFree(p->rays);
  if (p->order) (void)hipFree(p->order);
  if (p->tptr) (void)hipFree(p->tptr);
  if (p->tray) (void)hipFree(p->tray);
  if (p->tcol) (void)hipFree(p->tcol);
  if (p->scol) (void)hipFree(p->scol);
  if (p->sptr) (void)hipFree(p->sptr);
  if (p->srptr) (void)hipFree(p->srptr);
  if (p->sray) (void)hipFree(p->sray);
  if (p->sloc) (void)hipFree(p->sloc);
  if (p->cptr) (void)hipFree(p->cptr);
  if (p->cray) (void)hipFree(p->cray);
  delete p;
  return rc;
}

void tv3_plan_destroy(tv3_plan *p) {
  if (!p) return;
  // hipFree waits for the device; the plan's context may already be gone
  if (p->rays) (void)hipFree(p->rays);
  if (p->order) (void)hipFree(p->order);
  if (p->tptr) (void)hipFree(p->tptr);
  if (p->tray) (void)hipFree(p->tray);
  if (p->tcol) (void)hipFree(p->tcol);
  if (p->scol) (void)hipFree(p->scol);
  if (p->sptr) (void)hipFree(p->sptr);
  if (p->srptr) (void)hipFree(p->srptr);
  if (p->sray) (void)hipFree(p->sray);
  if (p->sloc) (void)hipFree(p->sloc);
  if (p->cptr) (void)hipFree(p->cptr);
  if (p->cray) (void)hipFree(p->cray);
  delete p;
}

int tv3_plan_shape(const tv3_plan *p, long long *m, long long *n,
                   long long *nnz) {
  if (!p) return TV3_ERR_ARG;
  if (m) *m = p->dev.m;
  if (n) *n = p->dev.n;
  if (nnz) *nnz = p->nnz;
  return TV3_OK;
}

int tv3_plan_walk(tv3_ctx *ctx, const tv3_plan *p, long long *ptr, int *lin) {
  if (!ctx || !p || !ptr || !lin) return TV3_ERR_ARG;
  if (p->dev.kind != kGeometry)
    return set_err(ctx, TV3_ERR_ARG, "Argument error: not a geometry plan");
  const long long m = p->dev.m, nnz = p->nnz;
  std::memcpy(ptr, p->walk_ptr.data(), sizeof(long long) * (m + 1));
  long long *d_ptr = nullptr;
  int *d_lin = nullptr;
  int rc = hip_check(ctx, hipMalloc(&d_ptr, sizeof(long long) * (m + 1)), "hipMalloc");
  if (!rc) rc = hip_check(ctx, hipMalloc(&d_lin, sizeof(int) * std::max<long long>(nnz, 1)), "hipMalloc");
  if (!rc) rc = hip_check(ctx, hipMemcpyAsync(d_ptr, ptr, sizeof(long long) * (m + 1), hipMemcpyHostToDevice, ctx->stream), "copy");
  if (!rc) {
    hipLaunchKernelGGL(k_walk_export, dim3((unsigned)((m + 3) / 4)), dim3(kBlock), 0, ctx->stream, p->dev.rays, (int)m, d_ptr, d_lin);
    rc = hip_check(ctx, hipGetLastError(), "k_walk_export");
  }
  if (!rc && nnz > 0) rc = hip_check(ctx, hipMemcpyAsync(lin, d_lin, sizeof(int) * nnz, hipMemcpyDeviceToHost, ctx->stream), "copy");
  if (!rc) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync");
  if (d_ptr) (void)hipFree(d_ptr);
  if (d_lin) (void)hipFree(d_lin);
  return rc;
}

}  // extern "C"
