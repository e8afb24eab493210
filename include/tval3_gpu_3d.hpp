// tval3_gpu_3d.hpp — tval3_gpu_3d: the reference's 3-D solver overload set
// (3d/tval3_cpu_3d.cpp:32-146 declarations, 1384-1580 definitions) on the GPU.
//
// Include after the reference's 3-D type header.  Same arguments, same
// argument checks and messages as tval3_cpu_3d (check_params,
// 3d/tval3_cpu_3d.cpp:1028-1165), same result record.  submat_mb sets
// the row blocks of the dense A^T A product (as in the reference);
// virtual_procs_per_core (MKL threads) is accepted and ignored.
// U's dim_z selects the 3-D stencil; dim_z == 1 is the 2-D one.  The Dt
// y-term of the reference (not the adjoint of D for dim_z >= 2) is kept.
//
// Supported element types: all of the reference's: float vectors with float
// or double scalars, double/double (through tv3_solve_f64) and
// complex<float>/float, complex<double>/double (through tv3_solve_c64 /
// _c128; Wz from the y direction as the reference computes it, 3d:934-937).
#ifndef TVAL3_GPU_3D_HPP_
#define TVAL3_GPU_3D_HPP_

#include "tval3_gpu_common.hpp"

namespace tval3_gpu_detail {

template <class Tc>
inline void check_3d(const mat<Tc> &U, const mat<Tc> &A, const mat<Tc> &b, const mat<Tc> &Ut) {
  if (U.format != mat_row_major || Ut.format != mat_row_major || A.format != mat_row_major)
    throw tval3_exception(std::string("Argument error: Matrices must be in row major format!"));
  else if (Ut.len > 0 && Ut.len != U.len)
    throw tval3_exception(std::string("Argument error: U and Ut must have the same size!"));
  else if (U.len != A.dim_x)
    throw tval3_exception(std::string("Argument error: the length of U must be equal to A.dim_x!"));
  else if (b.len != A.dim_y)
    throw tval3_exception(std::string("Argument error: b.len must be equal to A.dim_y!"));
}

template <class Tc>
inline void check_3d(const mat<Tc> &U, const sparse_mat<Tc> &A, const mat<Tc> &b,
                     const mat<Tc> &Ut) {
  if (U.format != mat_col_major || Ut.format != mat_col_major)
    throw tval3_exception(std::string("Argument error: Matrices must be in row major format!"));
  if (A.format != sparse_mat_both)
    throw tval3_exception(std::string("Argument error: A must be available in csr and csc format!"));
  if (Ut.len > 0 && Ut.len != U.len)
    throw tval3_exception(std::string("Argument error: U and Ut must have the same size!"));
  else if (U.len != A.dim_x)
    throw tval3_exception(std::string("Argument error: the length of U must be equal to A.dim_x!"));
  else if (b.len != A.dim_y)
    throw tval3_exception(std::string("Argument error: b.len must be equal to A.dim_y!"));
}

template <class Tc>
inline void check_3d(const mat<Tc> &U, const geometry &A, const mat<Tc> &b, const mat<Tc> &Ut) {
  if (U.format != mat_row_major || Ut.format != mat_row_major)
    throw tval3_exception(std::string("Argument error: Matrices must be in row major format!"));
  else if (Ut.len > 0 && Ut.len != U.len)
    throw tval3_exception(std::string("Argument error: U and Ut must have the same size!"));
  else if (b.len != A.numPaths * A.numMPs)
    throw tval3_exception(std::string("Argument error: b.len must be equal to A.num_emitters * A.num_receivers!"));
}

template <class Ts, class Tc>
inline tval3_info<Ts> run_3d(mat<Tc> &U, const mat<Tc> &A, const mat<Tc> &b,
                             const tval3_options<Ts> &opts, const mat<Tc> &Ut,
                             int submat_mb) {
  check_3d(U, A, b, Ut);
  return solve_host<tval3_exception, tval3_info<Ts>>(
      U, b, opts, Ut, (int)U.dim_x, (int)U.dim_y, (int)U.dim_z, sizeof(Ts) == 8, submat_mb,
      [&](tv3_ctx *c, PlanHolder &ph, KeepAll<tval3_exception> &keep) {
        dense_plan(c, ph, keep, &A.data()[0], (int)A.dim_y, (int)A.dim_x);
      });
}

template <class Ts, class Tc>
inline tval3_info<Ts> run_3d(mat<Tc> &U, const sparse_mat<Tc> &A, const mat<Tc> &b,
                             const tval3_options<Ts> &opts, const mat<Tc> &Ut,
                             int submat_mb) {
  check_3d(U, A, b, Ut);
  return solve_host<tval3_exception, tval3_info<Ts>>(
      U, b, opts, Ut, (int)U.dim_x, (int)U.dim_y, (int)U.dim_z, sizeof(Ts) == 8, submat_mb,
      [&](tv3_ctx *c, PlanHolder &ph, KeepAll<tval3_exception> &keep) {
        sparse_plan(c, ph, keep, A);
      });
}

template <class Ts, class Tc>
inline tval3_info<Ts> run_3d(mat<Tc> &U, const geometry &A, const mat<Tc> &b,
                             const tval3_options<Ts> &opts, const mat<Tc> &Ut,
                             int submat_mb) {
  check_3d(U, A, b, Ut);
  const int P = (int)A.numMPs, E = (int)A.num_emitters, R = (int)A.num_receivers;
  std::vector<int> up((size_t)E * R), xe((size_t)P * E), ye(xe.size()), ze(xe.size()),
      xr((size_t)P * R), yr(xr.size()), zr(xr.size());
  for (size_t k = 0; k < up.size(); ++k) up[k] = (int)A.use_path[k];
  for (int i = 0; i < P; ++i) {
    for (int e = 0; e < E; ++e) {
      xe[(size_t)i * E + e] = (int)A.x_emitters[i][e];
      ye[(size_t)i * E + e] = (int)A.y_emitters[i][e];
      ze[(size_t)i * E + e] = (int)A.z_emitters[i][e];
    }
    for (int r = 0; r < R; ++r) {
      xr[(size_t)i * R + r] = (int)A.x_receivers[i][r];
      yr[(size_t)i * R + r] = (int)A.y_receivers[i][r];
      zr[(size_t)i * R + r] = (int)A.z_receivers[i][r];
    }
  }
  tv3_geometry g{};
  g.num_mps = P;
  g.num_emitters = E;
  g.num_receivers = R;
  g.num_paths = (int)A.numPaths;
  g.use_path = up.data();
  g.x_emitters = xe.data();
  g.y_emitters = ye.data();
  g.z_emitters = ze.data();
  g.x_receivers = xr.data();
  g.y_receivers = yr.data();
  g.z_receivers = zr.data();
  g.rv_x = (int)A.rv_x;
  g.rv_y = (int)A.rv_y;
  g.rv_z = (int)A.rv_z;
  g.scale_factor = (float)A.scale_factor;
  return solve_host<tval3_exception, tval3_info<Ts>>(
      U, b, opts, Ut, (int)U.dim_x, (int)U.dim_y, (int)U.dim_z, sizeof(Ts) == 8, submat_mb,
      [&](tv3_ctx *c, PlanHolder &ph, KeepAll<tval3_exception> &) {
        geometry_plan<tval3_exception>(c, ph, g);
      });
}

}  // namespace tval3_gpu_detail

// ---- the overload set ---------------------------------------------------------
inline const tval3_info<float> tval3_gpu_3d(mat<float> &U, const mat<float> &A, const mat<float> &b,
                                            const tval3_options<float> &opts, const mat<float> &Ut,
                                            int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<float> tval3_gpu_3d(mat<float> &U, const sparse_mat<float> &A, const mat<float> &b,
                                            const tval3_options<float> &opts, const mat<float> &Ut,
                                            int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<float> tval3_gpu_3d(mat<float> &U, const geometry &A, const mat<float> &b,
                                            const tval3_options<float> &opts, const mat<float> &Ut,
                                            int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_3d(mat<float> &U, const mat<float> &A, const mat<float> &b,
                                             const tval3_options<double> &opts, const mat<float> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_3d(mat<float> &U, const sparse_mat<float> &A, const mat<float> &b,
                                             const tval3_options<double> &opts, const mat<float> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_3d(mat<float> &U, const geometry &A, const mat<float> &b,
                                             const tval3_options<double> &opts, const mat<float> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
// double / double
inline const tval3_info<double> tval3_gpu_3d(mat<double> &U, const mat<double> &A, const mat<double> &b,
                                             const tval3_options<double> &opts, const mat<double> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_3d(mat<double> &U, const sparse_mat<double> &A, const mat<double> &b,
                                             const tval3_options<double> &opts, const mat<double> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_3d(mat<double> &U, const geometry &A, const mat<double> &b,
                                             const tval3_options<double> &opts, const mat<double> &Ut,
                                             int submat_mb = 5, int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
// complex<float> / float, complex<double> / double
inline const tval3_info<float> tval3_gpu_3d(mat<std::complex<float> > &U, const mat<std::complex<float> > &A,
                                   const mat<std::complex<float> > &b, const tval3_options<float> &opts,
                                   const mat<std::complex<float> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<float> tval3_gpu_3d(mat<std::complex<float> > &U, const sparse_mat<std::complex<float> > &A,
                                   const mat<std::complex<float> > &b, const tval3_options<float> &opts,
                                   const mat<std::complex<float> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<float> tval3_gpu_3d(mat<std::complex<float> > &U, const geometry &A,
                                   const mat<std::complex<float> > &b, const tval3_options<float> &opts,
                                   const mat<std::complex<float> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_3d(mat<std::complex<double> > &U, const mat<std::complex<double> > &A,
                                   const mat<std::complex<double> > &b, const tval3_options<double> &opts,
                                   const mat<std::complex<double> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_3d(mat<std::complex<double> > &U, const sparse_mat<std::complex<double> > &A,
                                   const mat<std::complex<double> > &b, const tval3_options<double> &opts,
                                   const mat<std::complex<double> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}
inline const tval3_info<double> tval3_gpu_3d(mat<std::complex<double> > &U, const geometry &A,
                                   const mat<std::complex<double> > &b, const tval3_options<double> &opts,
                                   const mat<std::complex<double> > &Ut, int submat_mb = 5,
                                   int virtual_procs_per_core = 2) {
  (void)virtual_procs_per_core;
  return tval3_gpu_detail::run_3d(U, A, b, opts, Ut, submat_mb);
}

#endif  // TVAL3_GPU_3D_HPP_
