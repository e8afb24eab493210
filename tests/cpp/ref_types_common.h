// Minimal container types with the members the reference's solver uses
// (mat<T>: 2d/tval3_cpu_2d.cpp:157-195, 269-278; sparse_mat<T>:
// 2d/mat_mult.h:59-65; tval3_options<T>, tval3_info<T>: 2d:163-192, 865-876).
// The reference's own type header is not part of its tree; these exist only
// to compile the overload-layer tests.
#pragma once
#include <complex>
#include <exception>
#include <string>
#include <vector>

enum mat_format { mat_row_major, mat_col_major };
enum sparse_mat_format { sparse_mat_csr, sparse_mat_csc, sparse_mat_both };

template <class T>
struct mat {
  std::vector<T> buf;
  int dim_x = 0, dim_y = 0, dim_z = 1, len = 0;
  mat_format format = mat_row_major;
  mat() = default;
  explicit mat(int l) : buf(l), dim_x(1), dim_y(l), len(l) {}
  mat(int dy, int dx, int dz = 1) : buf((size_t)dx * dy * dz), dim_x(dx), dim_y(dy), dim_z(dz), len(dx * dy * dz) {}
  T *data() { return buf.data(); }
  const T *data() const { return buf.data(); }
  T &operator[](int i) { return buf[i]; }
};

template <class T>
struct sparse_mat {
  int dim_x = 0, dim_y = 0;
  sparse_mat_format format = sparse_mat_both;
  std::vector<int> csr_ptr, csr_ind, csc_ptr, csc_ind;
  std::vector<T> csr_val, csc_val;
};

template <class T>
struct tval3_options {
  T mu = 256, beta = 32, mu0 = 256, beta0 = 32, tol = 1e-4, tol_inn = 1e-3;
  int maxit = 300, maxcnt = 10;
  T gam = 0.9995, gamma = 0.6, rate_gam = 0.9, rate_cnt = 2, c = 1e-5;
  bool nonneg = false, isreal = true;
  unsigned numInitIterations = 0;
  T skipMulRatio = 0, maxRelativeChange = 0;
};

template <class T>
struct tval3_info {
  double secs = 0;
  int outer_iters = 0, total_iters = 0;
  T rel_chg = 0, rel_error = 0;
};

class tval3_exception : public std::exception {
 public:
  explicit tval3_exception(const std::string &m) : msg(m) {}
  explicit tval3_exception(const char *m) : msg(m) {}
  const char *what() const noexcept override { return msg.c_str(); }
 private:
  std::string msg;
};
