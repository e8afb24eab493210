/*
 * tval3_amd.h — C-ABI of the MI355X TVAL3 reconstruction engine.
 *
 * Plain C: POD structs, plain pointers and sizes, int status codes, no C++
 * exceptions and no torch types across the boundary.  Every entry point that
 * takes vector arguments takes DEVICE pointers and is asynchronous on the
 * context's HIP stream unless stated otherwise; geometry descriptions are
 * host arrays (they are turned into device-resident plans once).
 *
 * What each entry point replaces in the reference (modk/tval3-cpu):
 *   tv3_solve / tv3_solve_batch  <- tval3_cpu_2d / tval3_cpu_3d
 *                                   (2d/tval3_cpu_2d.cpp:752-886, 889-1016;
 *                                    3d/tval3_cpu_3d.cpp:1172-1378, 1384-1580)
 *   tv3_plan_dense               <- the mat<T> overloads of gemv
 *                                   (2d/mat_mult.h:15-36, 3d/mat_mult.h:17-45)
 *   tv3_plan_sparse              <- the sparse_mat<T> overloads of gemv
 *                                   (2d/mat_mult.h:38-91, 3d/mat_mult.h:47-118)
 *   tv3_plan_geometry            <- the geometry overload of gemv
 *                                   (3d/mat_mult.h:148-595; the 2-D copy at
 *                                    2d/mat_mult.h:107-194 is commented out)
 *   tv3_gemv                     <- gemv(CBLAS_TRANSPOSE, const T_A&, x, y)
 *   tv3_solve_f64,               <- the double/double overloads: mat<double>
 *   tv3_solve_batch_f64,            U with mat<double>, sparse_mat<double> or
 *   tv3_gemv_f64,                   geometry A (3d/tval3_cpu_3d.cpp:47, 88,
 *   tv3_plan_sparse_f64,            124; 2d/tval3_cpu_2d.cpp:26-29) and gemv
 *   tv3_plan_dense(TV3_F64)         on double vectors (3d/mat_mult.h:17-139)
 *   tv3_solve_c64 / _c128,       <- the complex overloads: mat<complex<float>>
 *   tv3_solve_batch_c64 / _c128,    / mat<complex<double>> U with complex
 *   tv3_gemv_c64 / _c128,           dense, complex sparse or geometry A
 *   tv3_plan_sparse_c64 / _c128,    (2d/tval3_cpu_2d.cpp:31-41, 58-68, 85-95;
 *   tv3_plan_dense(TV3_C64/C128)    3d/tval3_cpu_3d.cpp:56-65, 97-103,
 *                                   132-141), gemv with CblasConjTrans
 *   tv3_D / tv3_Dt               <- D / Dt (2d/tval3_cpu_2d.cpp:268-293,
 *                                   3d/tval3_cpu_3d.cpp:365-425)
 *   tv3_shrink                   <- shrinkage (2d/tval3_cpu_2d.cpp:588-611,
 *                                   3d/tval3_cpu_3d.cpp:834-871)
 *   tv3_dot / tv3_nrm2           <- cblas_dotc / cblas_nrm2 wrappers
 *                                   (2d/blas_wrap.h:71-112)
 *   tv3_geometry_save/load,      <- no counterpart: an on-disk form of the
 *   tv3_sparse_save/load            caller-filled geometry / sparse_mat<T>
 *   tv3_options / tv3_info       <- tval3_options<T> / tval3_info<T> (types
 *                                   from the reference's tval3_types.h, which
 *                                   the reference tree does not ship)
 *
 * Reduction semantics (both in this library and in the CPU oracle): every
 * dot product, norm and matrix-vector sum accumulates in FP64 and is rounded
 * once to the result type the reference uses (float for cblas_sdot/snrm2,
 * T_scal for the shrinkage sum).  "x + a*y" updates are single-rounding FMAs.
 */
#ifndef TVAL3_AMD_H_
#define TVAL3_AMD_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TV3_OK 0
#define TV3_ERR_ARG 1      /* argument error (reference: tval3_exception) */
#define TV3_ERR_HIP 2      /* HIP runtime error */
#define TV3_ERR_ALLOC 3    /* device allocation failed */
#define TV3_ERR_NOGPU 4    /* no HIP device */

/* ---- element types of a dense A ---------------------------------------- */
#define TV3_F32 0
#define TV3_BF16 1
#define TV3_F64 2 /* mat<double>: for tv3_solve_f64 / tv3_gemv_f64 only */
#define TV3_C64 3 /* mat<complex<float>>: for the _c64 entry points only */
#define TV3_C128 4 /* mat<complex<double>>: for the _c128 entry points only */

/* ---- gemv transpose flags (CBLAS_TRANSPOSE) ----------------------------- */
#define TV3_NOTRANS 0
#define TV3_TRANS 1        /* CblasTrans and CblasConjTrans (real data) */

/* Solver options: the 18 fields of the reference's tval3_options<T>
 * (used at 2d/tval3_cpu_2d.cpp:163-192, 540-569, 769-852), stored as double.
 * scal_double .. graph are engine settings with no field in tval3_options<T>;
 * submat_mb is the reference's submat_mb function argument. */
typedef struct tv3_options {
  double mu, beta;            /* final penalty parameters */
  double mu0, beta0;          /* initial penalty parameters (continuation) */
  double tol;                 /* outer stop: RelChgOut < tol */
  double tol_inn;             /* inner stop: RelChg < tol_inn */
  int maxit;                  /* cap on total inner iterations */
  int maxcnt;                 /* cap on outer iterations */
  double gam;                 /* nonmonotone weight */
  double gamma;               /* backtracking factor */
  double rate_gam;            /* gam shrink factor on a failed backtrack */
  double rate_cnt;            /* continuation rate for mu and beta */
  double c;                   /* Armijo constant */
  int nonneg;                 /* clamp U >= 0 after each descent step */
  int isreal;                 /* keep the real part (no-op for real data) */
  unsigned numInitIterations; /* A^T-skipping state machine */
  double skipMulRatio;        /* 0 disables A^T skipping */
  double maxRelativeChange;
  int scal_double;            /* 0: T_scal = float, 1: T_scal = double */
  int tv_iso;                 /* 0: anisotropic TV (reference), 1: isotropic */
  int dt_exact;               /* 3-D only. 0: the reference's Dt (its y-term
                                 is not the adjoint of D for dim_z >= 2) and,
                                 for complex data, its Wz from the y
                                 direction; 1: the exact adjoint of D and Wz
                                 from the z direction */
  int ls_exact;               /* Armijo backtracking.  1: the reference's
                                 rounds (each step re-interpolates every
                                 vector and re-evaluates f from rounded
                                 values); 0: f(alpha) of all five steps from
                                 the quadratic coefficients gathered by the
                                 descent pass, then one interpolation pass
                                 that recomputes the exact lam2..lam5 and f */
  int graph;                  /* 0: the host launches each step's kernels and
                                 skips branches no instance takes; 1: every
                                 step issues all (flag-masked) launches, two
                                 steps are captured into a hipGraph once and
                                 replayed, and the host only checks for
                                 termination.  Same results. */
  int submat_mb;              /* dense A: row-block size (MiB of float rows)
                                 of the blocked A^T A product of get_g, the
                                 reference's submat_mb argument (default 5;
                                 2d/tval3_cpu_2d.cpp:361-394) */
} tv3_options;

/* Result record: the reference's tval3_info<T> (2d/tval3_cpu_2d.cpp:865-876). */
typedef struct tv3_info {
  double secs;
  int outer_iters;
  int total_iters;
  double rel_chg;
  double rel_error;
} tv3_info;

/* One record per inner iteration (no reference counterpart: the reference
 * keeps no per-iteration history).  Written after the inner stop test. */
typedef struct tv3_trace_rec {
  int iter;        /* count_total after this iteration */
  int outer;       /* count_outer during this iteration */
  int backtracks;  /* Armijo backtracking steps taken (0..5) */
  int fallback;    /* 1 if the 5-strike fallback re-descended */
  double tau, alpha, f, C, rel_chg;
  double lam1, lam2, lam3, lam4, lam5;
  double mu, beta;
} tv3_trace_rec;

/* Ray geometry: the reference's `geometry` type as used by the live
 * projector (3d/mat_mult.h:164-232).  Host arrays.  Per-MP coordinate arrays
 * are flattened MP-major: x_emitters[mp * num_emitters + e].
 * use_path[e * num_receivers + r] = 0 skips the pair, k > 0 maps it to
 * output row mp * num_paths + k - 1. */
typedef struct tv3_geometry {
  int num_mps;
  int num_emitters;
  int num_receivers;
  int num_paths;
  const int *use_path;
  const int *x_emitters, *y_emitters, *z_emitters;
  const int *x_receivers, *y_receivers, *z_receivers;
  int rv_x, rv_y, rv_z;
  float scale_factor;
} tv3_geometry;

typedef struct tv3_ctx tv3_ctx;
typedef struct tv3_plan tv3_plan;

/* ---- defaults ------------------------------------------------------------ */
/* Fills the reference's option fields with the public TVAL3 defaults.  The
 * reference's own defaults live in a header it does not ship. */
void tv3_options_default(tv3_options *o);

/* ---- context --------------------------------------------------------------- */
/* stream: a hipStream_t, or NULL for a stream owned by the context. */
int tv3_ctx_create(int device, void *stream, tv3_ctx **out);
void tv3_ctx_destroy(tv3_ctx *ctx);
const char *tv3_ctx_error(const tv3_ctx *ctx);
int tv3_ctx_sync(tv3_ctx *ctx);
void *tv3_ctx_stream(const tv3_ctx *ctx);
/* Device memory for host callers (no reference counterpart: the reference
 * works on host arrays).  The C++ overload layer uses these to stage host
 * mat<T> data, so that it needs no HIP headers.  upload / download are
 * synchronous on the context stream. */
int tv3_mem_alloc(tv3_ctx *ctx, size_t bytes, void **out);
int tv3_mem_free(tv3_ctx *ctx, void *p);
int tv3_mem_upload(tv3_ctx *ctx, void *dst, const void *src, size_t bytes);
int tv3_mem_download(tv3_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Per-launch HIP-event timing of the solver's launches on the context stream
 * (it replaces the reference's PROFILING table, whose gemv slots 10/11 are
 * at 3d/mat_mult.h:19-44).  enable: 0 off, 1 the 64-lane or single-instance
 * A and A^T products (slots 10/11), 2 every launch (tv3_ctx_profile_slots),
 * 3 as 1 without the byte and lane counts of batched launches (no counting
 * kernel runs on the stream; profile_read then reports only the bytes of
 * launches the host counts); each call resets the totals.  profile_read returns, for [A, A^T], the summed
 * milliseconds, the number of launches and (bytes, may be NULL) the summed
 * compulsory bytes of those launches: the operator once for every group of
 * instances that computed (dense s_A*m*n; CSR / CSC (4 + s_v)*nnz +
 * 4(rows+1); geometry A x the 32-byte ray descriptors (and the 4-byte walk
 * order of the 64-lane kernel), A^T the 4*nnz + 8n + 4 bytes of the
 * voxel-major lists) plus s_vec*(n + m) of vectors for every instance that
 * computed.  profile_lanes returns, for [A, A^T], the number of instance
 * products those launches computed (a 64-lane launch with 40 computing
 * lanes counts 40). */
int tv3_ctx_profile(tv3_ctx *ctx, int enable);
int tv3_ctx_profile_read(tv3_ctx *ctx, double *ms, long long *count, double *bytes);
int tv3_ctx_profile_lanes(tv3_ctx *ctx, long long *lanes);

/* Per-slot timing report, the GPU counterpart of the reference's PROFILING
 * table (profile_info[slot]: name, time, runs; printed at 3d/tval3_cpu_3d.cpp:
 * 1324-1351).  With tv3_ctx_profile(ctx, 2) every launch of the solver is
 * timed by an event pair on the context stream (this slows a solve; level 1
 * times the two product slots only).  A slot number is the reference's
 * profile_info index where one pass implements that function (1 scaleb,
 * 3 shrinkage, 10 / 11 gemv NoTrans / ConjTrans), and 30 and up for fused
 * passes with no single counterpart (vec_add, nonneg, D and Dt, slots 0 and
 * 2 of the reference, run inside them; the names list what each contains;
 * 38 / 39 are the products of lanes gathered out of a 64-lane batch into
 * single-instance kernels, kept apart so that 10 / 11 time one kernel each).
 * Launches inside a captured graph (tv3_options.graph = 1) are not timed.
 * Sets *n to the number of slots that ran and fills out[0 .. min(cap, *n))
 * with them, in slot order (cap = 0 and out = NULL ask for the count);
 * synchronizes the stream. */
typedef struct tv3_profile_slot {
  int slot;
  long long runs;     /* launches */
  double ms;          /* summed milliseconds */
  char name[96];
} tv3_profile_slot;
int tv3_ctx_profile_slots(tv3_ctx *ctx, tv3_profile_slot *out, int cap, int *n);

/* ---- measurement-operator plans ------------------------------------------ */
/* Dense A, row-major m x n, on the device (f32 or bf16).  Borrowed. */
int tv3_plan_dense(tv3_ctx *ctx, int m, int n, int dtype, const void *A,
                   tv3_plan **out);
/* CSR and CSC copies of one m x n matrix, zero-based, device pointers,
 * borrowed (the reference requires format == sparse_mat_both). */
int tv3_plan_sparse(tv3_ctx *ctx, int m, int n, int nnz, const int *csr_ptr,
                    const int *csr_ind, const float *csr_val,
                    const int *csc_ptr, const int *csc_ind,
                    const float *csc_val, tv3_plan **out);
/* On-the-fly Bresenham projector.  Builds per-ray walk descriptors and the
 * deterministic voxel-major transpose once. */
int tv3_plan_geometry(tv3_ctx *ctx, const tv3_geometry *g, tv3_plan **out);
/* sparse_mat<double>: as tv3_plan_sparse with double values; for the _f64
 * entry points only (the float ones reject it, as they reject TV3_F64 dense
 * plans).  Geometry plans serve both. */
int tv3_plan_sparse_f64(tv3_ctx *ctx, int m, int n, int nnz, const int *csr_ptr,
                        const int *csr_ind, const double *csr_val,
                        const int *csc_ptr, const int *csc_ind,
                        const double *csc_val, tv3_plan **out);
/* sparse_mat<complex<float>> / sparse_mat<complex<double>>: values as
 * interleaved (re, im) pairs (2 * nnz components, aligned to the complex
 * element), for the _c64 / _c128 entry points only. */
int tv3_plan_sparse_c64(tv3_ctx *ctx, int m, int n, int nnz, const int *csr_ptr,
                        const int *csr_ind, const float *csr_val,
                        const int *csc_ptr, const int *csc_ind,
                        const float *csc_val, tv3_plan **out);
int tv3_plan_sparse_c128(tv3_ctx *ctx, int m, int n, int nnz, const int *csr_ptr,
                         const int *csr_ind, const double *csr_val,
                         const int *csc_ptr, const int *csc_ind,
                         const double *csc_val, tv3_plan **out);
/* Frees the plan's own device arrays; valid after its context is destroyed. */
void tv3_plan_destroy(tv3_plan *plan);
/* m = rows (measurements), n = columns (pixels), nnz = stored or walked
 * entries (m*n for dense). */
int tv3_plan_shape(const tv3_plan *plan, long long *m, long long *n,
                   long long *nnz);
/* Geometry plans: copies the voxel index list of every ray, rays in output
 * order, voxels in walk order, into host arrays ptr[m+1] and lin[nnz]. */
int tv3_plan_walk(tv3_ctx *ctx, const tv3_plan *plan, long long *ptr,
                  int *lin);

/* ---- operator files (host only; no GPU use; ctx may be NULL) -------------- */
/* The reference fills geometry and sparse_mat<T> in the caller
 * (3d/mat_mult.h:148-202, 2d/mat_mult.h:38-91) and has no file format; these
 * write and read one.  Little-endian and packed: "TV3GEOM1", int32 num_mps,
 * num_emitters, num_receivers, num_paths, rv_x, rv_y, rv_z, float32
 * scale_factor, int32 use_path[E*R], x/y/z_emitters[P*E], x/y/z_receivers[P*R];
 * "TV3SPAR1", int32 m, n, nnz, csr_ptr[m+1], csr_ind[nnz], float32
 * csr_val[nnz], int32 csc_ptr[n+1], csc_ind[nnz], float32 csc_val[nnz].
 * A load checks the magic, the exact length and every count and index, and
 * points the description into *storage, released with tv3_storage_free. */
int tv3_geometry_save(tv3_ctx *ctx, const char *path, const tv3_geometry *g);
int tv3_geometry_load(tv3_ctx *ctx, const char *path, tv3_geometry *g,
                      void **storage);
int tv3_sparse_save(tv3_ctx *ctx, const char *path, int m, int n, int nnz,
                    const int *csr_ptr, const int *csr_ind,
                    const float *csr_val, const int *csc_ptr,
                    const int *csc_ind, const float *csc_val);
int tv3_sparse_load(tv3_ctx *ctx, const char *path, int *m, int *n, int *nnz,
                    const int **csr_ptr, const int **csr_ind,
                    const float **csr_val, const int **csc_ptr,
                    const int **csc_ind, const float **csc_val,
                    void **storage);
void tv3_storage_free(void *storage);
/* Text of the calling thread's last failed operator-file call. */
const char *tv3_io_error(void);

/* ---- operators ------------------------------------------------------------ */
/* y = A x (trans = TV3_NOTRANS) or y = A^T x (TV3_TRANS); overwrites y. */
int tv3_gemv(tv3_ctx *ctx, const tv3_plan *A, int trans, const float *x,
             float *y);
/* tv3_gemv that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results as tv3_gemv; it calls no hipMalloc / hipFree, no
 * memset and no host copy.  Its group descriptor, instance record and, for
 * a dense A^T x, the 16 * n FP64 split-K partial sums live in one fixed
 * device array per GPU (1 KiB + 8 MiB, held until the process ends), so a
 * dense A^T x takes at most n = 65536 columns (TV3_ERR_ARG beyond that);
 * A x of any dense plan and sparse or geometry plans of any size take no
 * partials.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other threads, other contexts) race. */
int tv3_gemv_noalloc(tv3_ctx *ctx, const tv3_plan *A, int trans,
                     const float *x, float *y);
/* Periodic forward differences of U (dim_x contiguous).  Uz may be NULL
 * when dim_z == 1. */
int tv3_D(tv3_ctx *ctx, int dim_x, int dim_y, int dim_z, const float *U,
          float *Ux, float *Uy, float *Uz);
/* res = Dt(X, Y, Z); dt_exact as in tv3_options. */
int tv3_Dt(tv3_ctx *ctx, int dim_x, int dim_y, int dim_z, int dt_exact,
           const float *X, const float *Y, const float *Z, float *res);
/* W = shrink(U_dir - sigma_dir / beta, 1 / beta); returns lam1 (host). */
int tv3_shrink(tv3_ctx *ctx, int dim_x, int dim_y, int dim_z, double beta,
               int tv_iso, int scal_double, const float *Ux, const float *Uy,
               const float *Uz, const float *sx, const float *sy,
               const float *sz, float *Wx, float *Wy, float *Wz,
               double *lam1);
/* FP64-accumulated dot product and 2-norm (host results, synchronous). */
int tv3_dot(tv3_ctx *ctx, long long n, const float *x, const float *y,
            double *out);
int tv3_nrm2(tv3_ctx *ctx, long long n, const float *x, double *out);

/* ---- solver ---------------------------------------------------------------- */
/* One reconstruction (tval3_cpu_3d semantics; dim_z == 1 gives exactly the
 * tval3_cpu_2d semantics).  b: m floats, U: dim_x*dim_y*dim_z floats (output;
 * its input values are ignored, as in the reference), Ut: NULL or the ground
 * truth for info->rel_error.  trace: NULL or trace_cap host records.
 * Synchronous.
 * 2-D multi-layer mode (tval3_cpu_2d with b.len = layers * A.dim_y,
 * 2d/tval3_cpu_2d.cpp:889-1016, 2d/mat_mult.h:15-91): with dim_z == 1 and a
 * dense or sparse A, dim_x*dim_y may be `layers` times A's column count; U
 * then holds the layers' images one after the other, b holds layers * m
 * floats, and every product applies A to each layer. */
int tv3_solve(tv3_ctx *ctx, const tv3_plan *A, const tv3_options *opts,
              int dim_x, int dim_y, int dim_z, const float *b, float *U,
              const float *Ut, tv3_info *info, tv3_trace_rec *trace,
              int trace_cap);
/* B independent reconstructions advanced in lockstep by one set of batched
 * kernels.  A[i] may repeat (shared geometry); sizes may differ.  dims is
 * B x 3 (dim_x, dim_y, dim_z).  Ut may be NULL or contain NULL entries.
 * Synchronous. */
int tv3_solve_batch(tv3_ctx *ctx, int B, const tv3_plan *const *A,
                    const tv3_options *opts, const int *dims,
                    const float *const *b, float *const *U,
                    const float *const *Ut, tv3_info *infos);
/* The same with FP64 vectors and scalars (T_comp = T_scal = double;
 * opts->scal_double is ignored): A is a TV3_F64 dense, a double sparse or a
 * geometry plan, or a TV3_BF16 dense plan (no reference counterpart: its
 * entries widen exactly, so this is the FP64 solve of the bf16-rounded
 * matrix at the storage cost of bf16).  Every product, dot and norm is an FP64 sum, in a different
 * order than the reference's loops, so results agree with the reference's
 * double solve to FP64 rounding.  Instances run one per group (no 64-lane
 * interleaving); multi-layer mode as in tv3_solve. */
int tv3_solve_f64(tv3_ctx *ctx, const tv3_plan *A, const tv3_options *opts,
                  int dim_x, int dim_y, int dim_z, const double *b, double *U,
                  const double *Ut, tv3_info *info, tv3_trace_rec *trace,
                  int trace_cap);
int tv3_solve_batch_f64(tv3_ctx *ctx, int B, const tv3_plan *const *A,
                        const tv3_options *opts, const int *dims,
                        const double *const *b, double *const *U,
                        const double *const *Ut, tv3_info *infos);
/* y = A x or A^T x on double vectors (plans as for tv3_solve_f64). */
int tv3_gemv_f64(tv3_ctx *ctx, const tv3_plan *A, int trans, const double *x,
                 double *y);

/* The complex solves: vectors of interleaved (re, im) pairs (2 * len
 * components, aligned to the complex element), T_scal = the component type
 * (opts->scal_double has no effect; tv_iso = 1 is rejected: the reference
 * has no complex isotropic TV).  A is a TV3_C64 / TV3_C128 dense plan, a
 * tv3_plan_sparse_c64 / _c128 plan or a geometry plan (real weights).  Dots
 * are real(cblas_?dotc), shrinkage is per direction on |ubar| (2d:614-649),
 * nonneg clamps the real part and zeroes the imaginary one, opts->isreal = 1
 * keeps the real part of the result, and A^T is the conjugate transpose.
 * The reference's 3-D Wz takes the y direction's ubar (3d:934-937);
 * dt_exact = 1 takes the z direction's. */
int tv3_solve_c64(tv3_ctx *ctx, const tv3_plan *A, const tv3_options *opts,
                  int dim_x, int dim_y, int dim_z, const float *b, float *U,
                  const float *Ut, tv3_info *info, tv3_trace_rec *trace,
                  int trace_cap);
int tv3_solve_batch_c64(tv3_ctx *ctx, int B, const tv3_plan *const *A,
                        const tv3_options *opts, const int *dims,
                        const float *const *b, float *const *U,
                        const float *const *Ut, tv3_info *infos);
int tv3_gemv_c64(tv3_ctx *ctx, const tv3_plan *A, int trans, const float *x,
                 float *y);
int tv3_solve_c128(tv3_ctx *ctx, const tv3_plan *A, const tv3_options *opts,
                   int dim_x, int dim_y, int dim_z, const double *b, double *U,
                   const double *Ut, tv3_info *info, tv3_trace_rec *trace,
                   int trace_cap);
int tv3_solve_batch_c128(tv3_ctx *ctx, int B, const tv3_plan *const *A,
                         const tv3_options *opts, const int *dims,
                         const double *const *b, double *const *U,
                         const double *const *Ut, tv3_info *infos);
int tv3_gemv_c128(tv3_ctx *ctx, const tv3_plan *A, int trans, const double *x,
                  double *y);

/* ---- differentiation: K unrolled inner iterations and their adjoint ------- */
/* No single reference counterpart: the map is K inner iterations of the first
 * outer iteration of tval3_cpu_2d (2d/tval3_cpu_2d.cpp:752-830; min_u 526-576,
 * update_W 651-671, get_gradient 578-585) with no stopping test, sigma = delta
 * = 0, and the line-search decisions held fixed: alpha[k] is the step factor
 * of step k+1 and first_step[k] = 1 where that step used the first-step tau
 * (step 1, or a 5-strike fallback with alpha = 1).  A forward solve with
 * maxit = K, tol_inn = 0 and a trace gives both.  2-D, anisotropic TV,
 * without nonneg.  U = u_K / scl (device, dim_x * dim_y floats).  The tape
 * keeps checkpoints every ckpt steps; backward returns dL/db (device, m
 * floats) and dL/dmu (host) for the cotangent dU = dL/dU. */
typedef struct tv3_tape tv3_tape;
int tv3_unrolled_forward(tv3_ctx *ctx, const tv3_plan *A, const tv3_options *opts,
                         int dim_x, int dim_y, int K, int ckpt, const double *alpha,
                         const int *first_step, const float *b, float *U, tv3_tape **tape);
int tv3_unrolled_backward(tv3_ctx *ctx, tv3_tape *tape, const float *dU, float *db,
                          double *dmu);
void tv3_tape_destroy(tv3_tape *tape);

#ifdef __cplusplus
}
#endif
#endif /* TVAL3_AMD_H_ */
