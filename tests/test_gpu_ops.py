"""Operator-level parity of the HIP kernels against the CPU oracle.

Bit-exact: Bresenham voxel index lists, D, Dt (2-D, 3-D reference form and 3-D
exact form), shrinkage outputs.  Tolerance: matrix-vector products accumulate in
FP64 in a different order than the oracle and round once to float, so they may
differ by one float ulp; the tests bound the difference at 2 ulp-equivalents
(rtol 2.4e-7 of the row's absolute sum).
"""
import numpy as np
import pytest

from tval3_amd import synth
import tv3_oracle as O

pytestmark = pytest.mark.gpu


def _rand_geometry(seed, grid=(24, 20, 9), n_mp=2, E=7, R=9, drop=0.3):
    rng = np.random.default_rng(seed)
    gx, gy, gz = grid
    xe = rng.integers(0, gx, n_mp * E); ye = rng.integers(0, gy, n_mp * E)
    ze = rng.integers(0, gz, n_mp * E)
    xr = rng.integers(0, gx, n_mp * R); yr = rng.integers(0, gy, n_mp * R)
    zr = rng.integers(0, gz, n_mp * R)
    keep = rng.random(E * R) > drop
    up = np.zeros(E * R, np.int32)
    perm = rng.permutation(int(keep.sum()))           # rows not in loop order
    up[keep] = perm + 1
    npaths = int(keep.sum()) + 2                       # two rows no pair maps to
    return synth.GeometryArrays(n_mp, up, xe, ye, ze, xr, yr, zr, grid, npaths,
                                scale=0.75)


GEOS = {
    "ring64": lambda: synth.ring_geometry(64, 64),
    "ring128": lambda: synth.ring_geometry(128, 128),
    "hemi128_24": lambda: synth.hemisphere_geometry(128, 24),
    "rand3d": lambda: _rand_geometry(1),
    "rand3d_b": lambda: _rand_geometry(2, grid=(5, 33, 17), n_mp=3, E=11, R=4),
}


@pytest.mark.parametrize("name", list(GEOS))
def test_walk_bit_exact(ctx, name):
    geo = GEOS[name]()
    plan = ctx.plan_geometry(geo)
    op = O.Op.geometry(geo)
    assert plan.nnz == op.nnz
    gp, gl = plan.walk()
    op_, ol = op.walk()
    np.testing.assert_array_equal(gp, op_)
    np.testing.assert_array_equal(gl, ol)


def _close_rows(got, want, absrow):
    tol = 2.4e-7 * absrow + 1e-30
    bad = np.abs(got.astype(np.float64) - want) > tol
    assert not bad.any(), (np.flatnonzero(bad)[:5], got[bad][:5], want[bad][:5])


@pytest.mark.parametrize("name", list(GEOS))
def test_geometry_gemv(ctx, name):
    geo = GEOS[name]()
    plan = ctx.plan_geometry(geo)
    op = O.Op.geometry(geo)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(geo.n).astype(np.float32)
    y = rng.standard_normal(geo.m).astype(np.float32)
    ptr, lin = op.walk()
    # exact FP64 references from the walk
    rows = np.repeat(np.arange(geo.m), np.diff(ptr))
    sc = np.float32(geo.scale)
    fx = np.zeros(geo.m); np.add.at(fx, rows, (x[lin] * sc).astype(np.float64))
    fa = np.zeros(geo.m); np.add.at(fa, rows, np.abs((x[lin] * sc).astype(np.float64)))
    bx = np.zeros(geo.n); np.add.at(bx, lin, (y[rows] * sc).astype(np.float64))
    ba = np.zeros(geo.n); np.add.at(ba, lin, np.abs((y[rows] * sc).astype(np.float64)))
    g_f = ctx.gemv(plan, x).cpu().numpy()
    g_b = ctx.gemv(plan, y, trans=True).cpu().numpy()
    _close_rows(g_f, fx, fa)
    _close_rows(g_b, bx, ba)
    _close_rows(op.gemv(x), fx, fa)
    _close_rows(op.gemv(y, trans=True), bx, ba)
    # adjointness <Ax, y> = <x, A^T y>
    lhs = float(np.dot(g_f.astype(np.float64), y))
    rhs = float(np.dot(x.astype(np.float64), g_b))
    assert abs(lhs - rhs) <= 1e-5 * (np.abs(fa) @ np.abs(y) + 1)


@pytest.mark.parametrize("case", ["one_voxel", "no_ray"])
def test_geometry_mostly_uncrossed(ctx, case):
    """3-D volumes where (almost) no voxel is crossed: one ray whose emitter
    and receiver coincide (a walk of one voxel) and two rows no pair maps to,
    or no mapped pair at all.  The transpose writes the uncrossed voxels
    apart from the crossed ones (one thread each), so this covers a crossed
    part of one voxel and of none."""
    q = 16
    up = np.array([1 if case == "one_voxel" else 0], np.int32)
    geo = synth.GeometryArrays(1, up, [5], [7], [3], [5], [7], [3], (q, q, q), 3, scale=1.25)
    plan = ctx.plan_geometry(geo)
    op = O.Op.geometry(geo)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(geo.n).astype(np.float32)
    y = rng.standard_normal(geo.m).astype(np.float32)
    g_f = ctx.gemv(plan, x).cpu().numpy()
    g_b = ctx.gemv(plan, y, trans=True).cpu().numpy()
    np.testing.assert_array_equal(g_f, op.gemv(x))
    np.testing.assert_array_equal(g_b, op.gemv(y, trans=True))
    v = 5 + 7 * q + 3 * q * q
    want_b = np.zeros(geo.n, np.float32)
    want_f = np.zeros(geo.m, np.float32)
    if case == "one_voxel":
        want_b[v] = y[0] * np.float32(1.25)
        want_f[0] = x[v] * np.float32(1.25)
    np.testing.assert_array_equal(g_b, want_b)
    np.testing.assert_array_equal(g_f, want_f)


def test_sparse_gemv(ctx):
    geo = synth.ring_geometry(64, 64, scale=1.0)
    op = O.Op.geometry(geo)
    ptr, lin = op.walk()
    rng = np.random.default_rng(5)
    csr = list(synth.csr_from_walk(ptr, lin, geo.n))
    csr[2] = rng.uniform(0.5, 1.5, csr[2].size).astype(np.float32)
    # CSC values must follow the same matrix
    order = np.argsort(lin, kind="stable")
    csr[5] = csr[2][order]
    plan = ctx.plan_sparse(geo.m, geo.n, *csr)
    sop = O.Op.sparse(geo.m, geo.n, *csr)
    x = rng.standard_normal(geo.n).astype(np.float32)
    y = rng.standard_normal(geo.m).astype(np.float32)
    Ad = np.zeros((geo.m, geo.n))
    rows = np.repeat(np.arange(geo.m), np.diff(ptr))
    Ad[rows, lin] = csr[2]
    _close_rows(ctx.gemv(plan, x).cpu().numpy(), Ad @ x, np.abs(Ad) @ np.abs(x))
    _close_rows(ctx.gemv(plan, y, True).cpu().numpy(), Ad.T @ y, np.abs(Ad.T) @ np.abs(y))
    _close_rows(sop.gemv(x), Ad @ x, np.abs(Ad) @ np.abs(x))


@pytest.mark.parametrize("m,n", [(123, 257), (1229, 4096), (7, 5000)])
def test_dense_gemv(ctx, m, n):
    A = synth.gaussian_A(m, n, seed=m)
    plan = ctx.plan_dense(A)
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(np.float32)
    y = rng.standard_normal(m).astype(np.float32)
    Ad = A.astype(np.float64)
    _close_rows(ctx.gemv(plan, x).cpu().numpy(), Ad @ x, np.abs(Ad) @ np.abs(x))
    _close_rows(ctx.gemv(plan, y, True).cpu().numpy(), Ad.T @ y, np.abs(Ad.T) @ np.abs(y))


def test_dense_gemv_bf16(ctx):
    import torch
    m, n = 300, 1024
    A = synth.gaussian_A(m, n, seed=1)
    plan = ctx.plan_dense(A, bf16=True)
    Ab = torch.from_numpy(A).to(torch.bfloat16).to(torch.float64).numpy()
    x = np.random.default_rng(0).standard_normal(n).astype(np.float32)
    _close_rows(ctx.gemv(plan, x).cpu().numpy(), Ab @ x, np.abs(Ab) @ np.abs(x))


DIMS = [(17, 13, 1), (64, 64, 1), (9, 7, 5), (16, 12, 8), (5, 1, 3), (1, 6, 4)]


@pytest.mark.parametrize("dims", DIMS)
def test_D_bit_exact(ctx, dims):
    n = int(np.prod(dims))
    U = np.random.default_rng(n).standard_normal(n).astype(np.float32)
    got = ctx.D(U, dims)
    want = O.D(U, dims)
    for g, w in zip(got, want):
        if w is None:
            continue
        np.testing.assert_array_equal(g.cpu().numpy(), w)


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("exact", [False, True])
def test_Dt_bit_exact(ctx, dims, exact):
    n = int(np.prod(dims))
    rng = np.random.default_rng(n + 1)
    X, Y, Z = (rng.standard_normal(n).astype(np.float32) for _ in range(3))
    got = ctx.Dt(X, Y, Z, dims, exact).cpu().numpy()
    want = O.Dt(X, Y, Z, dims, exact)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("dims", [(31, 17, 1), (10, 9, 6)])
@pytest.mark.parametrize("iso", [False, True])
@pytest.mark.parametrize("sd", [False, True])
def test_shrink_bit_exact(ctx, dims, iso, sd):
    n = int(np.prod(dims))
    rng = np.random.default_rng(7)
    arrs = [rng.standard_normal(n).astype(np.float32) * 0.1 for _ in range(6)]
    if dims[2] == 1:
        arrs[2] = arrs[5] = None
    beta = 32.0 if not sd else 32.7
    Wx, Wy, Wz, lam = ctx.shrink(*arrs, dims, beta, iso, sd)
    oWx, oWy, oWz, olam = O.shrink(*arrs, dims, beta, iso, sd)
    np.testing.assert_array_equal(Wx.cpu().numpy(), oWx)
    np.testing.assert_array_equal(Wy.cpu().numpy(), oWy)
    if dims[2] > 1:
        np.testing.assert_array_equal(Wz.cpu().numpy(), oWz)
    assert lam == pytest.approx(olam, rel=2.4e-7 if not sd else 1e-14)


def test_dot(ctx):
    rng = np.random.default_rng(9)
    x = rng.standard_normal(100_003).astype(np.float32)
    y = rng.standard_normal(100_003).astype(np.float32)
    want = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
    assert ctx.dot(x, y) == pytest.approx(want, rel=1e-12, abs=1e-9)
    assert O.dot(x, y) == pytest.approx(want, rel=1e-12, abs=1e-9)
