// tv3_io.hip — on-disk form of the operator descriptions (host code only).
//
// The reference fills its `geometry` and `sparse_mat<T>` containers in the
// caller (3d/mat_mult.h:148-202 reads geometry.numMPs, num_emitters,
// num_receivers, numPaths, use_path, the per-position coordinate arrays,
// rv_x/y/z and scale_factor; 2d/mat_mult.h:38-91 reads the CSR and CSC
// arrays) and defines no file format.  These files let a geometry or sparse
// operator be written once and planned later.  Little-endian, packed:
//
//   geometry  "TV3GEOM1", int32 num_mps, num_emitters, num_receivers,
//             num_paths, rv_x, rv_y, rv_z, float32 scale_factor,
//             int32 use_path[E*R], x/y/z_emitters[P*E] (three arrays),
//             x/y/z_receivers[P*R] (three arrays)
//   sparse    "TV3SPAR1", int32 m, n, nnz, csr_ptr[m+1], csr_ind[nnz],
//             float32 csr_val[nnz], int32 csc_ptr[n+1], csc_ind[nnz],
//             float32 csc_val[nnz]
//
// Loading checks the magic, the exact file length, and every count and
// index a plan relies on; it reports the first violation as an argument
// error.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "tv3_internal.h"

using namespace tv3;

namespace {

const char kGeoMagic[8] = {'T', 'V', '3', 'G', 'E', 'O', 'M', '1'};
const char kSpMagic[8] = {'T', 'V', '3', 'S', 'P', 'A', 'R', '1'};

// arrays a loaded description points into
struct Storage {
  std::vector<int32_t> ints;
  std::vector<float> floats;
};

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

bool put(FILE *f, const void *p, size_t bytes) {
  return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes;
}

bool get(FILE *f, void *p, size_t bytes) {
  return bytes == 0 || std::fread(p, 1, bytes, f) == bytes;
}

// bytes left from the current position to the end of the file, or -1
long long remaining(FILE *f) {
  const long pos = std::ftell(f);
  if (pos < 0 || std::fseek(f, 0, SEEK_END) != 0) return -1;
  const long end = std::ftell(f);
  if (end < 0 || std::fseek(f, pos, SEEK_SET) != 0) return -1;
  return (long long)end - pos;
}

// message of the calling thread's last failed file call (also in ctx, if any)
thread_local std::string g_io_error;

int fail(tv3_ctx *ctx, int code, const std::string &msg) {
  g_io_error = msg;
  return set_err(ctx, code, msg);
}

int io_err(tv3_ctx *ctx, const std::string &path, const std::string &what) {
  return fail(ctx, TV3_ERR_ARG, "Argument error: " + path + ": " + what);
}

bool monotone(const int32_t *ptr, long long count, long long total) {
  if (ptr[0] != 0 || ptr[count] != total) return false;
  for (long long i = 0; i < count; ++i)
    if (ptr[i + 1] < ptr[i]) return false;
  return true;
}

bool in_range(const int32_t *ind, long long count, long long bound) {
  for (long long i = 0; i < count; ++i)
    if (ind[i] < 0 || ind[i] >= bound) return false;
  return true;
}

}  // namespace

extern "C" {

int tv3_geometry_save(tv3_ctx *ctx, const char *path, const tv3_geometry *g) {
  if (!path || !g) return fail(ctx, TV3_ERR_ARG, "Argument error: null argument");
  const long long P = g->num_mps, E = g->num_emitters, R = g->num_receivers;
  if (P <= 0 || E <= 0 || R <= 0 || g->num_paths <= 0 || !g->use_path || !g->x_emitters ||
      !g->y_emitters || !g->z_emitters || !g->x_receivers || !g->y_receivers || !g->z_receivers)
    return fail(ctx, TV3_ERR_ARG, "Argument error: invalid geometry");
  File fh;
  fh.f = std::fopen(path, "wb");
  if (!fh.f) return io_err(ctx, path, "cannot open for writing");
  const int32_t hdr[7] = {(int32_t)P, (int32_t)E, (int32_t)R, g->num_paths, g->rv_x, g->rv_y, g->rv_z};
  bool ok = put(fh.f, kGeoMagic, 8) && put(fh.f, hdr, sizeof(hdr)) &&
            put(fh.f, &g->scale_factor, 4) && put(fh.f, g->use_path, 4 * E * R);
  for (const int *a : {g->x_emitters, g->y_emitters, g->z_emitters}) ok = ok && put(fh.f, a, 4 * P * E);
  for (const int *a : {g->x_receivers, g->y_receivers, g->z_receivers}) ok = ok && put(fh.f, a, 4 * P * R);
  ok = ok && std::fflush(fh.f) == 0;
  return ok ? TV3_OK : io_err(ctx, path, "write failed");
}

int tv3_geometry_load(tv3_ctx *ctx, const char *path, tv3_geometry *g, void **storage) {
  if (!path || !g || !storage) return fail(ctx, TV3_ERR_ARG, "Argument error: null argument");
  *storage = nullptr;
  File fh;
  fh.f = std::fopen(path, "rb");
  if (!fh.f) return io_err(ctx, path, "cannot open");
  char magic[8];
  int32_t hdr[7];
  float scale;
  if (!get(fh.f, magic, 8) || std::memcmp(magic, kGeoMagic, 8) != 0)
    return io_err(ctx, path, "not a geometry file");
  if (!get(fh.f, hdr, sizeof(hdr)) || !get(fh.f, &scale, 4)) return io_err(ctx, path, "truncated header");
  const long long P = hdr[0], E = hdr[1], R = hdr[2], paths = hdr[3];
  if (P <= 0 || E <= 0 || R <= 0 || paths <= 0 || hdr[4] <= 0 || hdr[5] <= 0 || hdr[6] <= 0)
    return io_err(ctx, path, "invalid counts");
  const long long n_int = E * R + 3 * P * E + 3 * P * R;
  if (remaining(fh.f) != 4 * n_int) return io_err(ctx, path, "length does not match the counts");
  Storage *st = new (std::nothrow) Storage();
  if (!st) return fail(ctx, TV3_ERR_ALLOC, "out of host memory");
  try {
    st->ints.resize((size_t)n_int);
  } catch (...) {
    delete st;
    return fail(ctx, TV3_ERR_ALLOC, "out of host memory");
  }
  if (!get(fh.f, st->ints.data(), 4 * (size_t)n_int)) {
    delete st;
    return io_err(ctx, path, "read failed");
  }
  const int32_t *up = st->ints.data();
  for (long long k = 0; k < E * R; ++k)
    if (up[k] < 0 || up[k] > paths) {
      delete st;
      return io_err(ctx, path, "use_path entry out of range");
    }
  std::memset(g, 0, sizeof(*g));
  g->num_mps = (int)P;
  g->num_emitters = (int)E;
  g->num_receivers = (int)R;
  g->num_paths = (int)paths;
  g->rv_x = hdr[4];
  g->rv_y = hdr[5];
  g->rv_z = hdr[6];
  g->scale_factor = scale;
  const int32_t *p = up + E * R;
  g->use_path = up;
  g->x_emitters = p; p += P * E;
  g->y_emitters = p; p += P * E;
  g->z_emitters = p; p += P * E;
  g->x_receivers = p; p += P * R;
  g->y_receivers = p; p += P * R;
  g->z_receivers = p;
  *storage = st;
  return TV3_OK;
}

int tv3_sparse_save(tv3_ctx *ctx, const char *path, int m, int n, int nnz, const int *csr_ptr,
                    const int *csr_ind, const float *csr_val, const int *csc_ptr,
                    const int *csc_ind, const float *csc_val) {
  if (!path || !csr_ptr || !csc_ptr || (nnz > 0 && (!csr_ind || !csr_val || !csc_ind || !csc_val)))
    return fail(ctx, TV3_ERR_ARG, "Argument error: null argument");
  if (m <= 0 || n <= 0 || nnz < 0) return fail(ctx, TV3_ERR_ARG, "Argument error: invalid sparse matrix size");
  File fh;
  fh.f = std::fopen(path, "wb");
  if (!fh.f) return io_err(ctx, path, "cannot open for writing");
  const int32_t hdr[3] = {m, n, nnz};
  const size_t z = (size_t)nnz;
  const bool ok = put(fh.f, kSpMagic, 8) && put(fh.f, hdr, sizeof(hdr)) &&
                  put(fh.f, csr_ptr, 4 * ((size_t)m + 1)) && put(fh.f, csr_ind, 4 * z) &&
                  put(fh.f, csr_val, 4 * z) && put(fh.f, csc_ptr, 4 * ((size_t)n + 1)) &&
                  put(fh.f, csc_ind, 4 * z) && put(fh.f, csc_val, 4 * z) && std::fflush(fh.f) == 0;
  return ok ? TV3_OK : io_err(ctx, path, "write failed");
}

int tv3_sparse_load(tv3_ctx *ctx, const char *path, int *m, int *n, int *nnz, const int **csr_ptr,
                    const int **csr_ind, const float **csr_val, const int **csc_ptr,
                    const int **csc_ind, const float **csc_val, void **storage) {
  if (!path || !m || !n || !nnz || !csr_ptr || !csr_ind || !csr_val || !csc_ptr || !csc_ind ||
      !csc_val || !storage)
    return fail(ctx, TV3_ERR_ARG, "Argument error: null argument");
  *storage = nullptr;
  File fh;
  fh.f = std::fopen(path, "rb");
  if (!fh.f) return io_err(ctx, path, "cannot open");
  char magic[8];
  int32_t hdr[3];
  if (!get(fh.f, magic, 8) || std::memcmp(magic, kSpMagic, 8) != 0)
    return io_err(ctx, path, "not a sparse-matrix file");
  if (!get(fh.f, hdr, sizeof(hdr))) return io_err(ctx, path, "truncated header");
  const long long M = hdr[0], N = hdr[1], Z = hdr[2];
  if (M <= 0 || N <= 0 || Z < 0) return io_err(ctx, path, "invalid counts");
  if (remaining(fh.f) != 4 * ((M + 1) + (N + 1) + 4 * Z))
    return io_err(ctx, path, "length does not match the counts");
  Storage *st = new (std::nothrow) Storage();
  if (!st) return fail(ctx, TV3_ERR_ALLOC, "out of host memory");
  try {
    st->ints.resize((size_t)(M + 1 + N + 1 + 2 * Z));
    st->floats.resize((size_t)(2 * Z));
  } catch (...) {
    delete st;
    return fail(ctx, TV3_ERR_ALLOC, "out of host memory");
  }
  int32_t *rp = st->ints.data(), *ri = rp + (M + 1), *cp = ri + Z, *ci = cp + (N + 1);
  float *rv = st->floats.data(), *cv = rv + Z;
  const bool ok = get(fh.f, rp, 4 * (size_t)(M + 1)) && get(fh.f, ri, 4 * (size_t)Z) &&
                  get(fh.f, rv, 4 * (size_t)Z) && get(fh.f, cp, 4 * (size_t)(N + 1)) &&
                  get(fh.f, ci, 4 * (size_t)Z) && get(fh.f, cv, 4 * (size_t)Z);
  const char *bad = !ok ? "read failed"
                    : !monotone(rp, M, Z) || !monotone(cp, N, Z) ? "row or column pointers are not a partition of the entries"
                    : !in_range(ri, Z, N) ? "CSR column index out of range"
                    : !in_range(ci, Z, M) ? "CSC row index out of range"
                                          : nullptr;
  if (bad) {
    delete st;
    return io_err(ctx, path, bad);
  }
  *m = (int)M;
  *n = (int)N;
  *nnz = (int)Z;
  *csr_ptr = rp;
  *csr_ind = ri;
  *csr_val = rv;
  *csc_ptr = cp;
  *csc_ind = ci;
  *csc_val = cv;
  *storage = st;
  return TV3_OK;
}

void tv3_storage_free(void *storage) { delete static_cast<Storage *>(storage); }

const char *tv3_io_error(void) { return g_io_error.c_str(); }

}  // extern "C"
