"""Seeded synthetic inputs for the TVAL3 workloads (numpy only).

Every generator is a pure function of its seed, so tests, the smoke check and the
benchmark feed identical bytes to the GPU engine and to the CPU checker.

Workloads (sizes as used by the benchmark and the parity tests):
  dense 64x64        n=4096, m=1229 (30 %), A ~ N(0, 1/m) float32, 8-ellipse phantom
  ring 256x256       256 transducers on the inscribed circle, 32,640 pair rays
  hemisphere 128^3   1024 Fibonacci transducers on the lower hemisphere, 523,776 rays
  batched dense      grids from {64,96,128,192,256}^2, m ~ U[20 %, 40 %] of n
  ring 128x128       128 transducers, 8,128 rays (differentiation workload)
"""
import numpy as np

from .abi import Geometry
import ctypes as C

__all__ = [
    "gaussian_A", "ellipse_phantom", "ring_transducers", "ring_geometry",
    "disk_phantom", "hemisphere_transducers", "hemisphere_geometry",
    "ellipsoid_phantom", "add_noise", "GeometryArrays", "csr_from_walk",
    "batched_dense_sizes", "multi_mp_ring_geometry",
]


def gaussian_A(m, n, seed=0):
    """Dense A with i.i.d. N(0, 1/m) float32 entries, row-major m x n."""
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((m, n), dtype=np.float32)
            * np.float32(1.0 / np.sqrt(m))).astype(np.float32)


def ellipse_phantom(p, q, n_ell=8, seed=0):
    """Sum of n_ell random ellipses on a p x q grid (pixel centres in [-1, 1]^2):
    centres U[-0.6, 0.6]^2, semi-axes U[0.05, 0.4], angle U[0, pi),
    intensity U[0, 1].  Piecewise constant, float32, row-major (x contiguous)."""
    rng = np.random.default_rng(seed)
    ys = (np.arange(p) + 0.5) / p * 2 - 1
    xs = (np.arange(q) + 0.5) / q * 2 - 1
    X, Y = np.meshgrid(xs, ys)
    img = np.zeros((p, q), np.float64)
    for _ in range(n_ell):
        cx, cy = rng.uniform(-0.6, 0.6, 2)
        a, b = rng.uniform(0.05, 0.4, 2)
        th = rng.uniform(0, np.pi)
        val = rng.uniform(0, 1)
        xr = (X - cx) * np.cos(th) + (Y - cy) * np.sin(th)
        yr = -(X - cx) * np.sin(th) + (Y - cy) * np.cos(th)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += val
    return img.astype(np.float32).ravel()


class GeometryArrays:
    """Owns the int32 arrays behind a tv3_geometry and builds the ctypes struct."""

    def __init__(self, num_mps, use_path, xe, ye, ze, xr, yr, zr, rv,
                 num_paths, scale=1.0):
        self.num_mps = int(num_mps)
        self.use_path = np.ascontiguousarray(use_path, np.int32)
        self.xe, self.ye, self.ze = (np.ascontiguousarray(a, np.int32)
                                     for a in (xe, ye, ze))
        self.xr, self.yr, self.zr = (np.ascontiguousarray(a, np.int32)
                                     for a in (xr, yr, zr))
        self.rv = tuple(int(v) for v in rv)
        self.num_paths = int(num_paths)
        self.scale = float(scale)
        self.E = self.xe.size // self.num_mps
        self.R = self.xr.size // self.num_mps

    @property
    def m(self):
        return self.num_mps * self.num_paths

    @property
    def n(self):
        return self.rv[0] * self.rv[1] * self.rv[2]

    def struct(self):
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
        g = Geometry()
        g.num_mps, g.num_emitters, g.num_receivers = self.num_mps, self.E, self.R
        g.num_paths = self.num_paths
        g.use_path = ip(self.use_path)
        g.x_emitters, g.y_emitters, g.z_emitters = ip(self.xe), ip(self.ye), ip(self.ze)
        g.x_receivers, g.y_receivers, g.z_receivers = ip(self.xr), ip(self.yr), ip(self.zr)
        g.rv_x, g.rv_y, g.rv_z = self.rv
        g.scale_factor = self.scale
        return g


def _pairs_use_path(K):
    """use_path for all pairs i < j of K transducers (emitter-major rank, 1-based)."""
    up = np.zeros((K, K), np.int32)
    iu = np.triu_indices(K, 1)
    up[iu] = np.arange(1, iu[0].size + 1, dtype=np.int32)
    return up.ravel(), iu[0].size


def ring_transducers(K, grid):
    """K transducers equally spaced on the circle inscribed in a grid x grid image:
    (x, y) = rint(c + c*(cos t, sin t)), c = (grid-1)/2, t = 2 pi k / K."""
    c = (grid - 1) / 2.0
    t = 2 * np.pi * np.arange(K) / K
    return (np.rint(c + c * np.cos(t)).astype(np.int32),
            np.rint(c + c * np.sin(t)).astype(np.int32))


def ring_geometry(K=256, grid=256, scale=1.0):
    """2-D ultrasound ring: every emitter-receiver pair i < j (K(K-1)/2 rays) on a
    grid x grid image (rv_z = 1, z = 0)."""
    x, y = ring_transducers(K, grid)
    z = np.zeros(K, np.int32)
    up, npaths = _pairs_use_path(K)
    return GeometryArrays(1, up, x, y, z, x, y, z, (grid, grid, 1), npaths, scale)


def multi_mp_ring_geometry(K, grid, n_mps, skip_every=0, extra_rows=0, seed=0, scale=1.0):
    """Several measurement positions (numMPs > 1; 3d/mat_mult.h:183-202): at
    position i the ring of K transducers is rotated by 2 pi i / (K n_mps).
    use_path (shared by all positions) gives the pairs i < j a seeded random
    output order, 1-based, and leaves out every skip_every-th pair (0: none);
    row i * num_paths + use_path - 1 belongs to position i.  extra_rows rows
    per position get no pair (their measurements are 0)."""
    c = (grid - 1) / 2.0
    xs, ys = [], []
    for i in range(n_mps):
        t = 2 * np.pi * np.arange(K) / K + 2 * np.pi * i / (K * n_mps)
        xs.append(np.rint(c + c * np.cos(t)))
        ys.append(np.rint(c + c * np.sin(t)))
    x = np.concatenate(xs).astype(np.int32)
    y = np.concatenate(ys).astype(np.int32)
    z = np.zeros_like(x)
    iu = np.triu_indices(K, 1)
    keep = np.ones(iu[0].size, bool)
    if skip_every:
        keep[::skip_every] = False
    kept = int(keep.sum())
    num_paths = kept + int(extra_rows)
    rank = np.random.default_rng(seed).permutation(num_paths)[:kept].astype(np.int32) + 1
    up = np.zeros((K, K), np.int32)
    up[iu[0][keep], iu[1][keep]] = rank
    return GeometryArrays(n_mps, up.ravel(), x, y, z, x, y, z, (grid, grid, 1),
                          num_paths, scale)


def disk_phantom(grid=256, radius=None, seed=0):
    """Disk of value 1.0 with 5-10 circular inclusions (radius U[4, 20] px,
    contrast +-U[0.02, 0.1]) inside it.  float32, x contiguous."""
    rng = np.random.default_rng(seed)
    c = (grid - 1) / 2.0
    if radius is None:
        radius = grid * 120.0 / 256.0
    yy, xx = np.mgrid[0:grid, 0:grid].astype(np.float64)
    img = ((xx - c) ** 2 + (yy - c) ** 2 <= radius ** 2).astype(np.float64)
    k = rng.integers(5, 11)
    scale = grid / 256.0
    for _ in range(k):
        r = rng.uniform(4, 20) * scale
        rho = rng.uniform(0, radius - r - 1)
        ang = rng.uniform(0, 2 * np.pi)
        cx, cy = c + rho * np.cos(ang), c + rho * np.sin(ang)
        amp = rng.uniform(0.02, 0.1) * (1 if rng.random() < 0.5 else -1)
        img[(xx - cx) ** 2 + (yy - cy) ** 2 <= r * r] += amp
    return img.astype(np.float32).ravel()


def hemisphere_transducers(K=1024, grid=128):
    """K Fibonacci points on the lower hemisphere of radius (grid-1)/2 centred in
    the grid^3 cube, rounded to voxels."""
    c = (grid - 1) / 2.0
    k = np.arange(K)
    zf = -(k + 0.5) / K                      # z in (-1, 0): lower hemisphere
    r = np.sqrt(1 - zf * zf)
    phi = k * np.pi * (3 - np.sqrt(5))        # golden angle
    return (np.rint(c + c * r * np.cos(phi)).astype(np.int32),
            np.rint(c + c * r * np.sin(phi)).astype(np.int32),
            np.rint(c + c * zf).astype(np.int32))


def hemisphere_geometry(K=1024, grid=128, scale=1.0):
    """3-D ultrasound hemisphere: every pair i < j of K transducers."""
    x, y, z = hemisphere_transducers(K, grid)
    up, npaths = _pairs_use_path(K)
    return GeometryArrays(1, up, x, y, z, x, y, z, (grid, grid, grid), npaths,
                          scale)


def ellipsoid_phantom(grid=128, n_sph=20, seed=0):
    """Ellipsoid (semi-axes 60, 56, 52 voxels at grid 128, scaled) of value 1.0
    with n_sph spherical inclusions of +-0.05.  float32, x fastest."""
    rng = np.random.default_rng(seed)
    c = (grid - 1) / 2.0
    s = grid / 128.0
    ax = np.array([60, 56, 52]) * s
    zz, yy, xx = np.mgrid[0:grid, 0:grid, 0:grid].astype(np.float32)
    img = (((xx - c) / ax[0]) ** 2 + ((yy - c) / ax[1]) ** 2
           + ((zz - c) / ax[2]) ** 2 <= 1).astype(np.float32)
    for _ in range(n_sph):
        r = rng.uniform(3, 10) * s
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        rho = rng.uniform(0, 0.6) * ax.min()
        cx, cy, cz = c + rho * d
        amp = np.float32(0.05 if rng.random() < 0.5 else -0.05)
        img[(xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2 <= r * r] += amp
    return img.ravel()


def add_noise(y, rel, seed=0):
    """y + rel * rms(y) * N(0, 1), float32."""
    rng = np.random.default_rng(seed + 1_000_003)
    y = np.asarray(y, np.float64)
    rms = np.sqrt(np.mean(y * y)) if y.size else 0.0
    return (y + rel * rms * rng.standard_normal(y.size)).astype(np.float32)


def csr_from_walk(ptr, lin, n, scale=1.0):
    """CSR and CSC copies (zero-based int32, float32 values) of the matrix whose
    row r holds `scale` at the voxels of ray r's walk."""
    ptr = np.asarray(ptr, np.int64)
    lin = np.asarray(lin, np.int32)
    m = ptr.size - 1
    nnz = lin.size
    csr_ptr = ptr.astype(np.int32)
    csr_ind = lin.copy()
    csr_val = np.full(nnz, scale, np.float32)
    rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(ptr))
    order = np.argsort(lin, kind="stable")
    csc_ind = rows[order].astype(np.int32)
    csc_val = csr_val[order]
    counts = np.bincount(lin, minlength=n)
    csc_ptr = np.zeros(n + 1, np.int32)
    np.cumsum(counts, out=csc_ptr[1:])
    return (csr_ptr, csr_ind, csr_val, csc_ptr, csc_ind, csc_val)


def batched_dense_sizes(B=64, seed=0):
    """(grid, m) for B instances: grid from {64, 96, 128, 192, 256}, m ~ U[0.2, 0.4] n."""
    rng = np.random.default_rng(seed)
    grids = rng.choice([64, 96, 128, 192, 256], size=B)
    out = []
    for g in grids:
        n = int(g) * int(g)
        m = int(round(rng.uniform(0.2, 0.4) * n))
        out.append((int(g), m))
    return out
