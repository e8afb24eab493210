// Test driver of tval3_gpu_3d (include/tval3_gpu_3d.hpp).
//   shim_3d args                 argument checks only (no GPU use)
//   shim_3d dense|sparse|geometry DIR   as shim_2d, with dims dx dy dz
#include "ref_types_3d.h"
#include "tval3_gpu_3d.hpp"
#include "shim_io.h"

#include <algorithm>
#include <cstring>

static int args_mode() {
  mat<float> A(6, 64), b(6), U(4, 4, 4), Ut(4, 4, 4);
  tval3_options<float> of;
  mat<float> U_col(4, 4, 4);
  U_col.format = mat_col_major;
  expect_error("dense format", [&] { tval3_gpu_3d(U_col, A, b, of, Ut); });
  mat<float> Ut_bad(3, 3, 3);
  expect_error("dense Ut size", [&] { tval3_gpu_3d(U, A, b, of, Ut_bad); });
  mat<float> A_bad(6, 63);
  expect_error("dense U length", [&] { tval3_gpu_3d(U, A_bad, b, of, Ut); });
  mat<float> b_bad(7);
  expect_error("dense b length", [&] { tval3_gpu_3d(U, A, b_bad, of, Ut); });
  sparse_mat<float> S;
  S.dim_x = 64; S.dim_y = 6; S.format = sparse_mat_csc;
  mat<float> Uc(4, 4, 4), Utc(4, 4, 4);
  Uc.format = Utc.format = mat_col_major;
  expect_error("sparse format", [&] { tval3_gpu_3d(Uc, S, b, of, Utc); });
  geometry g;
  g.numMPs = 1; g.numPaths = 5;
  expect_error("geometry b length", [&] { tval3_gpu_3d(U, g, b, of, Ut); });
  // the double/double overloads run the same checks
  mat<double> Ud(4, 4, 4), Ad(6, 64), bd(7), Utd(4, 4, 4);
  tval3_options<double> od;
  expect_error("double b length", [&] { tval3_gpu_3d(Ud, Ad, bd, od, Utd); });
  // the complex overloads run the same checks
  mat<std::complex<float> > Ucx(4, 4, 4), Acx(6, 64), bcx(7), Utcx(4, 4, 4);
  expect_error("complex b length", [&] { tval3_gpu_3d(Ucx, Acx, bcx, of, Utcx); });
  return 0;
}

template <class Ts, class A_t>
static void run(const std::string &dir, const char *tag, const A_t &A, const mat<float> &b,
                int dx, int dy, int dz, bool col_major, int maxit) {
  mat<float> U(dy, dx, dz), Ut(0);
  U.format = Ut.format = col_major ? mat_col_major : mat_row_major;
  tval3_options<Ts> o;
  o.maxit = maxit;
  const tval3_info<Ts> info = tval3_gpu_3d(U, A, b, o, Ut);
  save(dir + "/U_" + tag + ".bin", U.data(), (size_t)U.len);
  save_info(dir + "/info_" + tag + ".txt", info);
}

static mat<double> to_double(const mat<float> &a) {
  mat<double> d(a.dim_y, a.dim_x, a.dim_z);
  d.format = a.format;
  std::copy(a.buf.begin(), a.buf.end(), d.buf.begin());
  return d;
}

static sparse_mat<double> to_double(const sparse_mat<float> &a) {
  sparse_mat<double> d;
  d.dim_x = a.dim_x; d.dim_y = a.dim_y; d.format = a.format;
  d.csr_ptr = a.csr_ptr; d.csr_ind = a.csr_ind; d.csc_ptr = a.csc_ptr; d.csc_ind = a.csc_ind;
  d.csr_val.assign(a.csr_val.begin(), a.csr_val.end());
  d.csc_val.assign(a.csc_val.begin(), a.csc_val.end());
  return d;
}

static const geometry &to_double(const geometry &g) { return g; }

// the double/double overload: U_dd.bin holds doubles
template <class A_t>
static void run_dd(const std::string &dir, const A_t &A, const mat<float> &b, int dx, int dy,
                   int dz, bool col_major, int maxit) {
  mat<double> U(dy, dx, dz), Ut(0);
  U.format = Ut.format = col_major ? mat_col_major : mat_row_major;
  const mat<double> bd = to_double(b);
  tval3_options<double> o;
  o.maxit = maxit;
  const tval3_info<double> info = tval3_gpu_3d(U, to_double(A), bd, o, Ut);
  save(dir + "/U_dd.bin", U.data(), (size_t)U.len);
  save_info(dir + "/info_dd.txt", info);
}

template <class R>
static mat<std::complex<R> > to_complex(const mat<float> &a) {
  mat<std::complex<R> > d(a.dim_y, a.dim_x, a.dim_z);
  d.format = a.format;
  for (size_t k = 0; k < a.buf.size(); ++k) d.buf[k] = std::complex<R>((R)a.buf[k], (R)0);
  return d;
}

template <class R>
static sparse_mat<std::complex<R> > to_complex(const sparse_mat<float> &a) {
  sparse_mat<std::complex<R> > d;
  d.dim_x = a.dim_x; d.dim_y = a.dim_y; d.format = a.format;
  d.csr_ptr = a.csr_ptr; d.csr_ind = a.csr_ind; d.csc_ptr = a.csc_ptr; d.csc_ind = a.csc_ind;
  for (float v : a.csr_val) d.csr_val.push_back(std::complex<R>((R)v, (R)0));
  for (float v : a.csc_val) d.csc_val.push_back(std::complex<R>((R)v, (R)0));
  return d;
}

template <class R>
static const geometry &to_complex(const geometry &g) { return g; }

// the complex<R>/R overload on b rotated by 0.6 + 0.8i (computed in double),
// imaginary part kept: U_c<tag>.bin holds (re, im) pairs
template <class R, class A_t>
static void run_cx(const std::string &dir, const char *tag, const A_t &A, const mat<float> &b,
                   int dx, int dy, int dz, bool col_major, int maxit) {
  mat<std::complex<R> > U(dy, dx, dz), Ut(0);
  U.format = Ut.format = col_major ? mat_col_major : mat_row_major;
  mat<std::complex<R> > bc(b.len);
  for (int k = 0; k < b.len; ++k)
    bc.buf[k] = std::complex<R>((R)(0.6 * (double)b.buf[k]), (R)(0.8 * (double)b.buf[k]));
  tval3_options<R> o;
  o.maxit = maxit;
  o.isreal = false;
  const tval3_info<R> info = tval3_gpu_3d(U, to_complex<R>(A), bc, o, Ut);
  save(dir + "/U_c" + tag + ".bin", U.data(), (size_t)U.len);
  save_info(dir + "/info_c" + tag + ".txt", info);
}

int main(int argc, char **argv) {
  try {
    if (argc >= 2 && !std::strcmp(argv[1], "args")) return args_mode();
    if (argc < 3) {
      std::fprintf(stderr, "usage: shim_3d args | dense|sparse|geometry DIR\n");
      return 2;
    }
    const std::string kind = argv[1], dir = argv[2];
    auto meta = load_meta(dir + "/meta.txt");
    const int dx = (int)meta["dx"], dy = (int)meta["dy"], dz = (int)meta["dz"];
    const int maxit = (int)meta["maxit"];
    std::vector<float> bv = load<float>(dir + "/b.bin");
    mat<float> b((int)bv.size());
    std::copy(bv.begin(), bv.end(), b.data());
    if (kind == "dense") {
      std::vector<float> Av = load<float>(dir + "/A.bin");
      mat<float> A((int)meta["m"], (int)meta["n"]);
      std::copy(Av.begin(), Av.end(), A.data());
      run<float>(dir, "f", A, b, dx, dy, dz, false, maxit);
      run<double>(dir, "d", A, b, dx, dy, dz, false, maxit);
      run_dd(dir, A, b, dx, dy, dz, false, maxit);
      run_cx<float>(dir, "f", A, b, dx, dy, dz, false, maxit);
      run_cx<double>(dir, "d", A, b, dx, dy, dz, false, maxit);
    } else if (kind == "sparse") {
      sparse_mat<float> S;
      S.dim_y = (int)meta["m"]; S.dim_x = (int)meta["n"];
      S.csr_ptr = load<int>(dir + "/csr_ptr.bin"); S.csr_ind = load<int>(dir + "/csr_ind.bin");
      S.csr_val = load<float>(dir + "/csr_val.bin"); S.csc_ptr = load<int>(dir + "/csc_ptr.bin");
      S.csc_ind = load<int>(dir + "/csc_ind.bin"); S.csc_val = load<float>(dir + "/csc_val.bin");
      run<float>(dir, "f", S, b, dx, dy, dz, true, maxit);
      run<double>(dir, "d", S, b, dx, dy, dz, true, maxit);
      run_dd(dir, S, b, dx, dy, dz, true, maxit);
      run_cx<float>(dir, "f", S, b, dx, dy, dz, true, maxit);
      run_cx<double>(dir, "d", S, b, dx, dy, dz, true, maxit);
    } else if (kind == "geometry") {
      const int P = (int)meta["mps"], E = (int)meta["E"], R = (int)meta["R"];
      std::vector<int> up = load<int>(dir + "/use_path.bin");
      std::vector<int> c[6] = {load<int>(dir + "/xe.bin"), load<int>(dir + "/ye.bin"),
                               load<int>(dir + "/ze.bin"), load<int>(dir + "/xr.bin"),
                               load<int>(dir + "/yr.bin"), load<int>(dir + "/zr.bin")};
      std::vector<int *> rows[6];
      for (int k = 0; k < 6; ++k) {
        const int w = k < 3 ? E : R;
        for (int i = 0; i < P; ++i) rows[k].push_back(c[k].data() + (size_t)i * w);
      }
      geometry g;
      g.numMPs = P; g.numPaths = (int)meta["paths"];
      g.num_emitters = E; g.num_receivers = R;
      g.use_path = up.data();
      g.x_emitters = rows[0].data(); g.y_emitters = rows[1].data(); g.z_emitters = rows[2].data();
      g.x_receivers = rows[3].data(); g.y_receivers = rows[4].data(); g.z_receivers = rows[5].data();
      g.rv_x = (int)meta["rv_x"]; g.rv_y = (int)meta["rv_y"]; g.rv_z = (int)meta["rv_z"];
      g.scale_factor = (float)meta["scale"];
      run<float>(dir, "f", g, b, dx, dy, dz, false, maxit);
      run<double>(dir, "d", g, b, dx, dy, dz, false, maxit);
      run_dd(dir, g, b, dx, dy, dz, false, maxit);
      run_cx<float>(dir, "f", g, b, dx, dy, dz, false, maxit);
      run_cx<double>(dir, "d", g, b, dx, dy, dz, false, maxit);
    } else {
      std::fprintf(stderr, "unknown kind %s\n", kind.c_str());
      return 2;
    }
    std::printf("ok\n");
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
