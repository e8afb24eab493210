// tval3_gpu_2d.hpp — tval3_gpu_2d: the reference's 2-D solver overload set
// (2d/tval3_cpu_2d.cpp:16-95 declarations, 889-1016 definitions) on the GPU.
//
// Include after the reference's 2-D type header.  Same arguments, same
// argument checks and messages as tval3_cpu_2d (check_params,
// 2d/tval3_cpu_2d.cpp:696-750), same result record.  submat_mb sets
// the row blocks of the dense A^T A product (as in the reference);
// virtual_procs_per_core (MKL threads) is accepted and ignored.
//
// Supported element types: all of the reference's: float vectors with float
// or double scalars (T_comp = float, T_scal = float | double), double/double
// (through tv3_solve_f64; results agree with the reference's double solve to
// FP64 rounding) and complex<float>/float, complex<double>/double (through
// tv3_solve_c64 / _c128).
// Multi-layer calls (b.len = layers * A.dim_y, U.len = layers * A.dim_x:
// U's pixels hold the layers one after the other, every product applies A
// to each layer, get_g takes two plain products) run for dense and sparse A,
// as in the reference.
//
// The reference's 2-D geometry projector is compiled out
// (2d/mat_mult.h:111-193: its gemv body is a comment, so A x = 0).  This
// overload runs the commented design instead: ray e * R + r for every pair
// with use_path != 0 (rows of skipped pairs stay 0), voxel x + y * ld,
// Bresenham steps of the 3-D projector with z = 0.
#ifndef TVAL3_GPU_2D_HPP_
#define TVAL3_GPU_2D_HPP_

#include "tval3_gpu_common.hpp"

namespace tval3_gpu_detail {

template <class Tc>
inline void check_2d(const mat<Tc> &U, const mat<Tc> &A, const mat<Tc> &b,
                     const mat<Tc> &Ut, int layers) {
  if (U.format != mat_col_major || Ut.format != mat_col_major || A.format != mat_row_major)
    throw tval3_exception(std::string("Argument error: Matrices must be in row major format!"));
  else if (Ut.len > 0 && Ut.len != U.len)
    throw tval3_exception(std::string("Argument error: U and Ut must have the same size!"));
  else if (U.len != A.dim_x * layers)
    throw tval3_exception(std::string("Argument error: the length of U must be equal to A.dim_x * layers!"));
  else if (b.len != A.dim_y * layers)
    throw tval3_exception(std::string("Argument error: b.len must be equal to A.dim_y * layers!"));
}

template <class Tc>
inline void check_2d(const mat<Tc> &U, const sparse_mat<Tc> &A, const mat<Tc> &b,
                     const mat<Tc> &Ut, int layers) {
  if (U.format != mat_col_major || Ut.format != mat_col_major)
    throw tval3_exception(std::string("Argument error: Matrices must be in row major format!"));
  if (A.format != sparse_mat_both)
    throw tval3_exception(std::string("Argument error: A must be available in CSR and CSC format!"));
  if (Ut.len > 0 && Ut.len != U.len)
    throw tval3_exception(std::string("Argument error: U and Ut must have the same size!"));
  else if (U.len != A.dim_x * layers)
    throw tval3_exception(std::string("Argument error: the length of U must be equal to A.dim_x * layers!"));
  else if (b.len != A.dim_y * layers)
    throw tval3_exception(std::string("Argument error: b.len must be equal to A.dim_y * layers!"));
}

template <class Tc>
inline void check_2d(const mat<Tc> &U, const geometry &A, const mat<Tc> &b,
                     const mat<Tc> &Ut, int layers) {
  if (U.format != mat_row_major || Ut.format != mat_row_major)
    throw tval3_exception(std::string("Argument error: Matrices must be in row major format!"));
  else if (Ut.len > 0 && Ut.len != U.len)
    throw tval3_exception(std::string("Argument error: U and Ut must have the same size!"));
  else if (layers > 1)
    throw tval3_exception(std::string("Argument error: multilayer reconstruction with dynamic calculation is not supported!"));
  else if (b.len != A.num_emitters * A.num_receivers)
    throw tval3_exception(std::string("Argument error: b.len must be equal to A.num_emitters * A.num_receivers!"));
}

template <class Ts, class Tc>
inline tval3_info<Ts> run_2d(mat<Tc> &U, const mat<Tc> &A, const mat<Tc> &b,
                             const tval3_options<Ts> &opts, const mat<Tc> &Ut,
                             int submat_mb) {
  const int layers = A.dim_y ? (int)(b.len / A.dim_y) : 0;
  check_2d(U, A, b, Ut, layers);
  return solve_host<tval3_exception, tval3_info<Ts>>(
      U, b, opts, Ut, (int)U.dim_x, (int)U.dim_y, 1, sizeof(Ts) == 8, submat_mb,
      [&](tv3_ctx *c, PlanHolder &ph, KeepAll<tval3_exception> &keep) {
        dense_plan(c, ph, keep, &A.data()[0], (int)A.dim_y, (int)A.dim_x);
      });
}

template <class Ts, class Tc>
inline tval3_info<Ts> run_2d(mat<Tc> &U, const sparse_mat<Tc> &A, const mat<Tc> &b,
                             const tval3_options<Ts> &opts, const mat<Tc> &Ut,
                             int submat_mb) {
  const int layers = A.dim_y ? (int)(b.len / A.dim_y) : 0;
  check_2d(U, A, b, Ut, layers);
  return solve_host<tval3_exception, tval3_info<Ts>>(
      U, b, opts, Ut, (int)U.dim_x, (int)U.dim_y, 1, sizeof(Ts) == 8, submat_mb,
      [&](tv3_ctx *c, PlanHolder &ph, KeepAll<tval3_exception> &keep) {
        sparse_plan(c, ph, keep, A);
      });
}

template <class Ts, class Tc>
inline tval3_info<Ts> run_2d(mat<Tc> &U, const geometry &A, const mat<Tc> &b,
                             const tval3_options<Ts> &opts, const mat<Tc> &Ut,
                             int submat_mb) {
  check_2d(U, A, b, Ut, 1);
  const int E = (int)A.num_emitters, R = (int)A.num_receivers;
  std::vector<int> up((size_t)E * R), xe(E), ye(E), ze(E, 0), xr(R), yr(R), zr(R, 0);
  for (int e = 0; e < E; ++e) {
    xe[e] = (int)A.x_emitters[e];
    ye[e] = (int)A.y_emitters[e];
    for (int r = 0; r < R; ++r) up[(size_t)e * R + r] = A.use_path[(size_t)e * R + r] ? e * R + r + 1 : 0;
  }
  for (int r = 0; r < R; ++r) {
    xr[r] = (int)A.x_receivers[r];
    yr[r] = (int)A.y_receivers[r];
  }
  tv3_geometry g{};
  g.num_mps = 1;
  g.num_emitters = E;
  g.num_receivers = R;
  g.num_paths = E * R;
  g.use_path = up.data();
  g.x_emitters = xe.data();
  g.y_emitters = ye.data();
  g.z_emitters = ze.data();
  g.x_receivers = xr.data();
  g.y_receivers = yr.data();
  g.z_receivers = zr.data();
  g.rv_x = (int)A.ld;
  g.rv_y = A.ld ? (int)(U.len / A.ld) : 0;
  g.rv_z = 1;
  g.scale_factor = (float)A.scale_factor;
  return solve_host<tval3_exception, tval3_info<Ts>>(
      U, b, opts, Ut, (int)U.dim_x, (int)U.dim_y, 1, sizeof(Ts) == 8, submat_mb,
      [&](tv3_ctx *c, PlanHolder &ph, KeepAll<tval3_exception> &) {
        geometry_plan<tval3_exception>(c, ph, g);
      });
}

}  // namespace tval3_gpu_detail

// ---- the overload set ---------------------------------------------------------
// float / float
inline const tval3_info<float> tval3_gpu_2d(mat<float> &U, const mat<float> &A, const mat<float> &b,
                                            const tval3_options<float> &opts, const mat<float> &Ut,
                                            int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<float> tval3_gpu_2d(mat<float> &U, const sparse_mat<float> &A, const mat<float> &b,
                                            const tval3_options<float> &opts, const mat<float> &Ut,
                                            int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<float> tval3_gpu_2d(mat<float> &U, const geometry &A, const mat<float> &b,
                                            const tval3_options<float> &opts, const mat<float> &Ut,
                                            int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
// float / double
inline const tval3_info<double> tval3_gpu_2d(mat<float> &U, const mat<float> &A, const mat<float> &b,
                                             const tval3_options<double> &opts, const mat<float> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_2d(mat<float> &U, const sparse_mat<float> &A, const mat<float> &b,
                                             const tval3_options<double> &opts, const mat<float> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_2d(mat<float> &U, const geometry &A, const mat<float> &b,
                                             const tval3_options<double> &opts, const mat<float> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
// double / double
inline const tval3_info<double> tval3_gpu_2d(mat<double> &U, const mat<double> &A, const mat<double> &b,
                                             const tval3_options<double> &opts, const mat<double> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_2d(mat<double> &U, const sparse_mat<double> &A, const mat<double> &b,
                                             const tval3_options<double> &opts, const mat<double> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_2d(mat<double> &U, const geometry &A, const mat<double> &b,
                                             const tval3_options<double> &opts, const mat<double> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
// complex<float> / float, complex<double> / double
inline const tval3_info<float> tval3_gpu_2d(mat<std::complex<float> > &U, const mat<std::complex<float> > &A,
                                   const mat<std::complex<float> > &b, const tval3_options<float> &opts,
                                   const mat<std::complex<float> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<float> tval3_gpu_2d(mat<std::complex<float> > &U, const sparse_mat<std::complex<float> > &A,
                                   const mat<std::complex<float> > &b, const tval3_options<float> &opts,
                                   const mat<std::complex<float> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<float> tval3_gpu_2d(mat<std::complex<float> > &U, const geometry &A,
                                   const mat<std::complex<float> > &b, const tval3_options<float> &opts,
                                   const mat<std::complex<float> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_2d(mat<std::complex<double> > &U, const mat<std::complex<double> > &A,
                                   const mat<std::complex<double> > &b, const tval3_options<double> &opts,
                                   const mat<std::complex<double> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_2d(mat<std::complex<double> > &U, const sparse_mat<std::complex<double> > &A,
                                   const mat<std::complex<double> > &b, const tval3_options<double> &opts,
                                   const mat<std::complex<double> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_2d(mat<std::complex<double> > &U, const geometry &A,
                                   const mat<std::complex<double> > &b, const tval3_options<double> &opts,
                                   const mat<std::complex<double> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_2d(U, A, b, opts, Ut, submat_mb);
}

#endif  // TVAL3_GPU_2D_HPP_
