// tv3_internal.h — shared host/device declarations of the MI355X TVAL3 engine.
//
// Memory layout.  A batch of reconstructions is split into groups of W lanes
// (W = 1 for a single or uneven-size batch, W = 64 for instances that share a
// measurement operator and image size).  Every vector of a group is stored
// lane-interleaved: element v of lane b lives at base + v * W + b, so a
// wavefront that walks one ray for 64 instances issues 256-byte coalesced
// loads, and a wavefront with W = 1 walks 64 consecutive pixels.  Per-instance
// scalars (penalties, Lagrangian terms, step sizes, counters, control flags)
// stay on the device in InstScal; the host reads back only a flag summary.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/tval3_amd.h"

namespace tv3 {

constexpr int kBlock = 256;       // threads per block of every vector kernel
constexpr int kNblkMax = 2048;    // partial-sum entries per slot: max blocks of a
                                  // single-instance (W = 1) reduction kernel
constexpr int kNblkGroup = 512;   // max blocks per group of a 16/64-lane reduction
                                  // kernel, and of the stand-alone operations
constexpr int kNslot = 24;        // partial-sum slots per instance
constexpr int kSlotM = 12;        // first slot written by m-vector kernels
// Im <s, y> of the complex Barzilai-Borwein step: written by the same pass as
// <s, s> (slot 8) and Re <s, y> (slot 9), read by the next get_tau; no kernel
// between the two writes slot 10
constexpr int kSlotSyIm = 10;

// ---- vector element types ----------------------------------------------------
// Complex elements of the complex solvers (the reference's complex<float> /
// complex<double> T_comp): x = real part, y = imaginary part.  Only the
// operations the solver needs are defined: sums and differences, negation and
// scaling by a real value, each per component.  A real value converts to
// (value, 0).
struct __align__(8) cfloat {
  float x, y;
  cfloat() = default;
  __host__ __device__ constexpr cfloat(float r) : x(r), y(0.0f) {}
  __host__ __device__ constexpr cfloat(float r, float i) : x(r), y(i) {}
};
struct __align__(16) cdouble {
  double x, y;
  cdouble() = default;
  __host__ __device__ constexpr cdouble(double r) : x(r), y(0.0) {}
  __host__ __device__ constexpr cdouble(double r, double i) : x(r), y(i) {}
};
#define TV3_CPLX_OPS(C, R)                                                                  \
  __host__ __device__ inline C operator+(C a, C b) { return C(a.x + b.x, a.y + b.y); }    \
  __host__ __device__ inline C operator-(C a, C b) { return C(a.x - b.x, a.y - b.y); }    \
  __host__ __device__ inline C operator-(C a) { return C(-a.x, -a.y); }                   \
  __host__ __device__ inline C operator*(R s, C a) { return C(s * a.x, s * a.y); }         \
  __host__ __device__ inline C operator*(C a, R s) { return C(a.x * s, a.y * s); }
TV3_CPLX_OPS(cfloat, float)
TV3_CPLX_OPS(cdouble, double)
#undef TV3_CPLX_OPS

// component type of a vector element
template <class T>
struct vreal {
  typedef T type;
  static constexpr bool cplx = false;
};
template <>
struct vreal<cfloat> {
  typedef float type;
  static constexpr bool cplx = true;
};
template <>
struct vreal<cdouble> {
  typedef double type;
  static constexpr bool cplx = true;
};

// ---- plans ------------------------------------------------------------------
enum PlanKind { kDense = 0, kSparse = 1, kGeometry = 2 };

// One output row of the on-the-fly projector.  Voxel j (0 <= j <= L) of the
// walk is lin0 + j*s_major + ka(j)*s_a + kb(j)*s_b with
// k(j) = floor((2*m*j + L - 1) / (2*L)), the closed form of the Bresenham
// error recurrence of the reference (3d/mat_mult.h:241-359).  L < 0 marks a
// row no pair maps to (its value is 0).
struct RayDesc {
  int lin0;
  int L;
  int s_major, s_a, s_b;
  int ma, mb;
  float inv2L;  // 1 / (2L) for the division fix-up
};

// Device view of a plan, embedded in every group descriptor.
struct PlanDev {
  int kind;
  int m, n;
  // dense
  const void *A;
  int dtype;
  int a_vec;  // A is 16-byte aligned and n % 8 == 0: rows take vector loads
  long long nnz;  // stored entries (dense: m*n; geometry: walk length)
  // sparse (CSR of A and CSR of A^T = CSC of A)
  const int *r_ptr, *r_ind;
  const float *r_val;
  const int *c_ptr, *c_ind;
  const float *c_val;
  // double values of a sparse_mat<double> plan (then r_val / c_val are null)
  const double *r_vald, *c_vald;
  // complex values of a sparse_mat<complex<T>> plan: cfloat (cval = 2) or
  // cdouble (cval = 3) elements; 0: none
  const void *r_valx, *c_valx;
  int cval;
  // geometry
  const RayDesc *rays;
  const int *g_cptr;  // voxel-major transpose: rays through voxel v,
  const int *g_cray;  // ascending, at g_cray[g_cptr[v] .. g_cptr[v+1])
  // order in which the 64-lane forward walks the rays: rays between
  // neighbouring emitters and neighbouring receivers come together
  const int *ray_order;
  // the transpose lists again in the 64-lane transpose's walk order: slot s
  // is voxel t_col[s], its rays t_ray[t_ptr[s] .. t_ptr[s+1]); the voxels no
  // ray crosses come last, from slot t_nz on
  const int *t_ptr, *t_ray, *t_col;
  int t_nz;
  // the staged transpose of the single-instance kernels (k_bwd_w1s): slot
  // s < s_nz is voxel s_col[s] (bricks of the volume; lists of at most
  // kStageH entries), the s_nh voxels with longer lists follow (longest
  // first), then the voxels no ray crosses.  Tile t is slots
  // [t * s_V, (t + 1) * s_V);
  // its rays s_ray[s_rptr[t] .. s_rptr[t + 1]) (ascending) are staged in
  // LDS.  Slot s's list is s_loc[s_ptr[s] .. s_ptr[s + 1]): the voxel's
  // g_cray list in order, each ray given by its position among the tile's
  // rays.  A tile crossed by more than kStageR rays stages none
  // (s_rptr[t] == s_rptr[t + 1]) and reads g_cray.  s_col null: no staging.
  const int *s_col, *s_ptr, *s_rptr, *s_ray;
  const unsigned short *s_loc;
  int s_nz, s_nh, s_V;
  long long s_rays;  // s_rptr[tiles]
  long long s_dnnz;  // list entries read as ray ids: the longer lists and
                     // the tiles that stage no rays
  float scale;
};

// Vector types a plan serves: float-valued operators (f32 / bf16 dense, float
// sparse) multiply float vectors, double-valued ones (f64 dense, double
// sparse) double vectors; the geometry projector serves both.
__host__ __device__ inline bool plan_takes_f32(const PlanDev &p) {
  return p.kind == kDense ? (p.dtype == TV3_F32 || p.dtype == TV3_BF16)
                          : p.kind == kSparse ? (p.r_val != nullptr || (p.nnz == 0 && p.cval == 0)) : true;
}
// (a bf16 dense A also serves double vectors: widened exactly, it gives the
// FP64 solve of the bf16-rounded matrix)
__host__ __device__ inline bool plan_takes_f64(const PlanDev &p) {
  return p.kind == kDense ? (p.dtype == TV3_F64 || p.dtype == TV3_BF16)
                          : p.kind == kSparse ? p.r_vald != nullptr || p.nnz == 0 : true;
}
// complex vectors: complex entries of the same precision, or the projector
__host__ __device__ inline bool plan_takes_c64(const PlanDev &p) {
  return p.kind == kDense ? p.dtype == TV3_C64 : p.kind == kSparse ? p.cval == 2 : true;
}
__host__ __device__ inline bool plan_takes_c128(const PlanDev &p) {
  return p.kind == kDense ? p.dtype == TV3_C128 : p.kind == kSparse ? p.cval == 3 : true;
}

// ---- batch description ------------------------------------------------------
struct GroupDesc {
  int q, p, r;          // dim_x, dim_y, dim_z
  int n, m;             // pixels, measurements (plan.n, plan.m times layers)
  // 2-D multi-layer mode: the n-vectors hold `layers` images of plan.n
  // pixels one after the other and the m-vectors `layers` blocks of plan.m;
  // the operator applies A to each (W = 1 batches only)
  int layers;
  long long off_n;      // float offset of this group's n-vectors
  long long off_m;      // float offset of this group's m-vectors
  PlanDev plan;
  // blocked dense product of get_g (W = 1): rows per block, slices per block
  // (0: one thread walks all rows of its columns) and the double offset of
  // this group's slice partials in the solver scratch
  int dblk_rows, dblk_S;
  long long dblk_off;
};

// Control flags (InstScal::flags)
enum : uint32_t {
  F_ACTIVE = 1u << 0,     // instance still iterating
  F_FST = 1u << 1,        // first inner iteration of an outer iteration
  F_BT = 1u << 2,         // Armijo failed: backtracking round pending
  F_FB = 1u << 3,         // 5-strike fallback pending
  F_BWD = 1u << 4,        // A^T product needed in this descent
  F_INNER_DONE = 1u << 5, // inner loop ended this step
  F_MLP = 1u << 6,        // multiplier update pending
  F_JUST_DONE = 1u << 7,  // finished in this step (write the output)
  F_DONE = 1u << 8,       // finished
  F_SKIPG = 1u << 9,      // A^T skipped: g = gp - Atb
  F_DESC = 1u << 10,      // descent pending (set for all active at a step)
  F_LIVE = 1u << 11,      // lane holds a real instance (not padding)
};

// Per-instance scalars.  T_scal quantities are kept in double; for the
// float instantiation every value is a float widened exactly.  f, C, lam4,
// lam5 are T_comp in the reference and are stored in the vector type VC
// (float, or double for the FP64-vector solver) in a 32-byte base that
// comes first, so both instantiations have the same size and every other
// field (flags included) the same offset: the operator launches, which read
// only `flags`, take either through InstScal.
template <class VC>
struct InstScalVC;
template <>
struct InstScalVC<float> {
  float lam4, lam5, f, C;
  float pad_vc[4];
};
template <>
struct InstScalVC<double> {
  double lam4, lam5, f, C;
};
template <class VC>
struct InstScalT : InstScalVC<VC> {
  uint32_t flags;
  int count_outer, count_total, cnt;  // cnt: backtracking steps this iter
  int fallback;
  unsigned currInitIteration;
  double mu, beta, muDbeta, beta0;
  double lam1, lam2, lam3;
  double gam, Q;
  double tau, alpha, c_armij;
  double RelChg, RelChgOut, nrmup;
  double scl;
  double currSkipValue;
  double ratio;        // beta0 / beta after a continuation step
  double rel_error;
  int trace_n;
  int pad;
};
typedef InstScalT<float> InstScal;
typedef InstScalT<double> InstScalD;
static_assert(sizeof(InstScalVC<float>) == sizeof(InstScalVC<double>) &&
                  sizeof(InstScal) == sizeof(InstScalD),
              "instance scalar layouts differ");

// Options converted once to device form.
struct OptDev {
  double mu, beta, mu0, beta0, tol, tol_inn;
  int maxit, maxcnt;
  double gam, gamma, rate_gam, rate_cnt, c;
  int nonneg, tv_iso, dt_exact, skip, ls_exact;
  unsigned numInitIterations;
  double skipMulRatio, maxRelativeChange;
};

}  // namespace tv3

// ---- context and plan objects (host) ------------------------------------------
struct tv3_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  // grow-only device workspace for the solver
  void *ws = nullptr;
  size_t ws_bytes = 0;
  void *pinned = nullptr;  // small pinned host buffer
  size_t pinned_bytes = 0;
  // per-instance control flags, written by the device straight into mapped
  // pinned host memory (read after a stream synchronization)
  // published instance flags in mapped host memory (host and device views):
  // one 64-bit word per instance, the flags in the low half and the number
  // of the launch that wrote them in the high half; pub_epoch is the last
  // number given out
  unsigned long long *hflags = nullptr, *dflags = nullptr;
  size_t hflags_n = 0;
  uint32_t pub_epoch = 0;
  // timing events for the benchmark (kernel-level, on this stream)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // optional per-launch timing (tv3_ctx_profile): level 1 times the A / A^T
  // products, level 2 every engine launch, into the slots of tv3::kProf*
  bool prof = false;
  int prof_level = 0;
  bool prof_open = false;      // a begin event of the current launch is recorded
  std::vector<hipEvent_t> evpool;
  std::vector<int> evkind;     // slot of each recorded pair
  double prof_ms[40] = {};     // per slot (tv3::kProfSlots)
  long long prof_n[40] = {};
  // device: algorithmic bytes of the instances each profiled launch computed
  unsigned long long *prof_bytes = nullptr;
  // host: the same for launches whose one instance the host knows computes
  unsigned long long prof_hbytes[4] = {0, 0, 0, 0};  // bytes A, A^T; lanes A, A^T
};

struct tv3_plan {
  tv3_ctx *ctx = nullptr;
  tv3::PlanDev dev{};
  long long nnz = 0;
  // owned device arrays (geometry)
  tv3::RayDesc *rays = nullptr;
  int *cptr = nullptr, *cray = nullptr, *order = nullptr;
  int *tptr = nullptr, *tray = nullptr, *tcol = nullptr;
  int *scol = nullptr, *sptr = nullptr, *srptr = nullptr, *sray = nullptr;
  unsigned short *sloc = nullptr;
  std::vector<long long> walk_ptr;  // host: start of each row's walk
};

namespace tv3 {
// error helpers (tv3_plan.hip)
int set_err(tv3_ctx *ctx, int code, const std::string &msg);
// Profile slots: the reference's profile_info indices where one pass
// implements that reference function (3d/tval3_cpu_3d.cpp:255, 314, 837;
// 3d/mat_mult.h:21-27), 30 and up for fused passes with no single
// counterpart.
constexpr int kProfSlots = 40;
static_assert(sizeof(tv3_ctx::prof_ms) == kProfSlots * sizeof(double), "slot arrays");
constexpr int kProfScaleb = 1, kProfShrink = 3, kProfGemvN = 10, kProfGemvT = 11;
constexpr int kProfDescend = 30, kProfLineSearch = 31, kProfTau = 32, kProfGetG = 33,
              kProfMlp = 34, kProfScalars = 35, kProfLanes = 36, kProfSetup = 37,
              kProfGemvNLanes = 38, kProfGemvTLanes = 39;
const char *prof_slot_name(int slot);
// event pair around one launch when profiling is on (level 1: the A / A^T
// slots only)
void prof_begin(tv3_ctx *ctx, int kind);
void prof_end(tv3_ctx *ctx);
void prof_collect(tv3_ctx *ctx);
// dense W = 1 launch shape: rows per wavefront of A x (1, 2, 4), rows
// loaded ahead in A^T x (1, 2, 4, 8) and columns per thread of the blocked
// A^T x (4, 8).  Neither changes any summation order.
// The current setting is per host thread; the solver sets it from its
// autotuner for the duration of a solve.
struct DenseTune {
  int rows;
  int unroll;  // 1, 2, 4, 8
  int cols;    // columns per thread of the blocked A^T product: 2, 4 or 8
  // no group of the batch runs the whole-row blocked kernel (1) or the
  // row-slice kernels (1): their launches are skipped
  int no_blk = 0, no_slice = 0;
  // multi-layer batch (every group holds layers >= 2 of one shared A): the
  // products run as FP64-MFMA GEMMs over the layers (k_dense_*_mfma).  The
  // largest layer count, A row and column counts of the batch size the
  // grids; vec_f: every group's A takes vector row loads (a_vec), vec_b:
  // every group's m is a multiple of 4 (vector loads of the layers'
  // m-vectors)
  int mfma = 0, mfma_layers = 0, mfma_m = 0, mfma_n = 0, mfma_vec_f = 0, mfma_vec_b = 0;
};
void dense_tune_set(DenseTune t);
DenseTune dense_tune_cur();

// adds the algorithmic bytes of the lanes a masked A (trans 0) / A^T (trans 1)
// launch computed to ctx->prof_bytes[trans]; W lanes per group, Bpad lanes
// (host-known launches: `groups` groups of width W read the operator,
// `lanes` lanes computed)
void prof_count_host(tv3_ctx *ctx, const GroupDesc &g, int trans, int vbytes, int W = 1,
                     int groups = 1, int lanes = 1);
void prof_count(tv3_ctx *ctx, const GroupDesc *G_dev, int W, int Bpad,
                const InstScal *S, uint32_t mask, int trans, int vbytes);
int hip_check(tv3_ctx *ctx, hipError_t e, const char *what);

// operator launches (tv3_plan.hip).  x and y are batch base pointers;
// flag_mask selects the lanes that compute (0 = all).
// kind and dtype are those of every group of the launch (one operator kind
// per batch; dense batches of mixed element type are split by the caller).
template <int W>
void launch_fwd(hipStream_t s, const GroupDesc *G_dev, int G, int max_m,
                int kind, int dtype, const float *x, float *y,
                const InstScal *S, uint32_t mask);
// submat_mb > 0 (dense A only): the blocked product of get_g, row blocks of
// max(submat_mb MiB / 4 / n, 1) rows, each block's sum rounded to float and
// the blocks added in float in order; 0: one FP64 sum over all rows.
template <int W>
void launch_bwd(hipStream_t s, const GroupDesc *G_dev, int G, int max_n,
                int kind, int dtype, const float *x, float *y,
                const InstScal *S, uint32_t mask, const float *sub,
                float *scratch, int submat_mb);

// FP64 vectors (the double/double solver): W = 1 only.  submat_mb is
// accepted and ignored: the transposed product is one split FP64 sum.
template <int W>
void launch_fwd(hipStream_t s, const GroupDesc *G_dev, int G, int max_m,
                int kind, int dtype, const double *x, double *y,
                const InstScal *S, uint32_t mask);
template <int W>
void launch_bwd(hipStream_t s, const GroupDesc *G_dev, int G, int max_n,
                int kind, int dtype, const double *x, double *y,
                const InstScal *S, uint32_t mask, const double *sub,
                double *scratch, int submat_mb);

// Complex vectors (the complex solvers): W = 1 only.  submat_mb > 0 (dense
// A): the blocked A^H product of get_g, block sums rounded to the component
// type and added per component in block order.
template <int W>
void launch_fwd(hipStream_t s, const GroupDesc *G_dev, int G, int max_m,
                int kind, int dtype, const cfloat *x, cfloat *y,
                const InstScal *S, uint32_t mask);
template <int W>
void launch_bwd(hipStream_t s, const GroupDesc *G_dev, int G, int max_n,
                int kind, int dtype, const cfloat *x, cfloat *y,
                const InstScal *S, uint32_t mask, const cfloat *sub,
                cfloat *scratch, int submat_mb);
template <int W>
void launch_fwd(hipStream_t s, const GroupDesc *G_dev, int G, int max_m,
                int kind, int dtype, const cdouble *x, cdouble *y,
                const InstScal *S, uint32_t mask);
template <int W>
void launch_bwd(hipStream_t s, const GroupDesc *G_dev, int G, int max_n,
                int kind, int dtype, const cdouble *x, cdouble *y,
                const InstScal *S, uint32_t mask, const cdouble *sub,
                cdouble *scratch, int submat_mb);

// chunks of the split-K dense A^T x (plain products: Atb, tv3_gemv): its
// scratch holds kDenseChunks * n FP64 partial sums per group
constexpr int kDenseChunks = 16;

// The no-allocation gemv (tv3_gemv_noalloc, tv3_solver.hip) keeps its group
// descriptor, instance record and dense partial sums in one file-scope
// device array per GPU; noalloc_attach looks up that array's address on
// `device` (current) once and keeps it.  tv3_ctx_create calls it.
int noalloc_attach(int device);

// rows per block of the blocked dense product (2d/tval3_cpu_2d.cpp:366-369):
// submat_mb MiB of elem-byte entries
__host__ __device__ inline int dense_block_rows(int submat_mb, int n, int elem = 4) {
  const long long e = (long long)submat_mb * 1024 * 1024 / elem;
  const long long r = e / (n > 0 ? n : 1);
  return r < 1 ? 1 : (int)(r > 0x7fffffff ? 0x7fffffff : r);
}
}  // namespace tv3
