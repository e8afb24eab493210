"""The complex solvers (tv3_solve_c64 / _c128, tv3_gemv_c64 / _c128) against
the oracle's complex instantiations.

complex<float> shares every elementwise rounding with the oracle (per
component, the complex magnitude of the shrinkage in float as the reference
computes it) and sums in FP64 in another order, like the float path: the
reconstruction is held to 1e-4 relative L2 with identical trajectories under
ls_exact = 1.  complex<double> is held to the FP64 bar of f64_bar (10x the
oracle's own spread under other reduction orders).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import tval3_amd
from tval3_amd import synth
import tv3_oracle as O
import f64_bar as F

pytestmark = pytest.mark.gpu

TOL_U = 1e-4


def _crand(rng, shape, t=np.complex64):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(t)


def _compare(ctx, plan, op, b, dims, opts, strict_traj=True, cap=400):
    U, info, tr = ctx.solve(plan, b, dims, opts, trace_cap=cap)
    Ug = U.cpu().numpy()
    Uo, info_o, tro = O.solve_c(op, b, dims, opts, trace_cap=cap)
    rel = F.rel(Ug, Uo)
    diff = [t.iter for t, s in zip(tr, tro) if (t.backtracks, t.fallback) != (s.backtracks, s.fallback)]
    wide = b.dtype == np.complex128
    bar = F.bound(op, b, dims, opts, Uo, O.solve_c) if wide else TOL_U
    msg = (f"rel={rel:.3e} bar={bar:.3e} exact={np.array_equal(Ug, Uo)} "
           f"iters gpu/oracle={info.total_iters}/{info_o.total_iters} first traj diff at {diff[:3]}")
    print(msg)
    assert Ug.dtype == b.dtype
    assert np.all(np.isfinite(Ug)), msg
    assert rel <= bar, msg
    if strict_traj:
        assert info.total_iters == info_o.total_iters, msg
        assert info.outer_iters == info_o.outer_iters, msg
        assert not diff, msg
    return Ug, Uo


def _dense(seed, m=150, n=256, t=np.complex64):
    rng = np.random.default_rng(seed)
    A = (synth.gaussian_A(m, n, seed) + 1j * synth.gaussian_A(m, n, seed + 100)).astype(t)
    x = synth.ellipse_phantom(16, 16, 4, seed).ravel() * np.exp(0.3j)
    b = (A.astype(np.complex128) @ x + 0.01 * _crand(rng, m, np.complex128)).astype(t)
    return A, b


@pytest.mark.parametrize("t,tol", [(np.complex64, 1e-6), (np.complex128, 1e-13)])
def test_complex_gemv(ctx, t, tol):
    """Dense and sparse complex A against numpy (A^H is the conjugate
    transpose); the projector against the oracle per component."""
    rng = np.random.default_rng(7)
    for m, n in ((37, 53), (300, 700)):
        A = _crand(rng, (m, n), t)
        x, y = _crand(rng, n, t), _crand(rng, m, t)
        Ad = A.astype(np.complex128)
        plan = ctx.plan_dense(A)
        assert F.rel(ctx.gemv(plan, x).cpu().numpy(), Ad @ x) <= tol
        assert F.rel(ctx.gemv(plan, y, trans=True).cpu().numpy(), Ad.conj().T @ y) <= tol
        As = A * (rng.random((m, n)) < 0.2)
        S, T = sp.csr_matrix(As), sp.csc_matrix(As)
        splan = ctx.plan_sparse(m, n, S.indptr, S.indices, S.data, T.indptr, T.indices, T.data)
        Asd = As.astype(np.complex128)
        assert F.rel(ctx.gemv(splan, x).cpu().numpy(), Asd @ x) <= tol
        assert F.rel(ctx.gemv(splan, y, trans=True).cpu().numpy(), Asd.conj().T @ y) <= tol
    for geo in (synth.ring_geometry(48, 48), synth.hemisphere_geometry(40, 12)):
        op = O.Op.geometry(geo)
        plan = ctx.plan_geometry(geo)
        x, y = _crand(rng, geo.n, t), _crand(rng, geo.m, t)
        gx = ctx.gemv(plan, x).cpu().numpy()
        gy = ctx.gemv(plan, y, trans=True).cpu().numpy()
        if t == np.complex64:
            # the walk sums of the float path, per component
            np.testing.assert_array_equal(gx, op.gemv_c(x))
            np.testing.assert_array_equal(gy, op.gemv_c(y, trans=True))
        else:
            assert F.rel(gx, op.gemv_c(x)) <= tol and F.rel(gy, op.gemv_c(y, trans=True)) <= tol


@pytest.mark.parametrize("ls", [1, 0])
def test_dense_complex_solve(ctx, ls):
    A, b = _dense(11)
    opts = tval3_amd.default_options(maxit=80, ls_exact=ls, isreal=0)
    _compare(ctx, ctx.plan_dense(A), O.Op.dense_c(A), b, (16, 16, 1), opts, strict_traj=bool(ls))


def test_dense_complex_blocked_and_layers(ctx):
    """The blocked A^H A of get_g with several row blocks (submat_mb small),
    and the multi-layer mode (plain products)."""
    A, b = _dense(12, m=400, n=256)
    opts = tval3_amd.default_options(maxit=60, ls_exact=1, isreal=0, submat_mb=1)
    # 1 MiB blocks of 8-byte entries over 256 columns: 512 rows per block, so
    # A has one block and A2 three
    A2, b2 = _dense(13, m=1200, n=256)
    for AA, bb in ((A, b), (A2, b2)):
        _compare(ctx, ctx.plan_dense(AA), O.Op.dense_c(AA), bb, (16, 16, 1), opts)
    L = 3
    x = np.concatenate([synth.ellipse_phantom(16, 16, 4, 20 + l).ravel() for l in range(L)])
    bl = np.concatenate([(A.astype(np.complex128) @ x[l * 256:(l + 1) * 256]) * (1 + 0.2j)
                         for l in range(L)]).astype(np.complex64)
    _compare(ctx, ctx.plan_dense(A), O.Op.dense_c(A), bl, (16, 16 * L, 1), opts)


def test_sparse_complex_and_options(ctx):
    rng = np.random.default_rng(14)
    A, b = _dense(14)
    As = A * (rng.random(A.shape) < 0.3)
    S, T = sp.csr_matrix(As), sp.csc_matrix(As)
    args = (As.shape[0], As.shape[1], S.indptr, S.indices, S.data, T.indptr, T.indices, T.data)
    plan, op = ctx.plan_sparse(*args), O.Op.sparse_c(*args)
    for kw in ({}, {"nonneg": 1}, {"isreal": 1}):
        opts = tval3_amd.default_options(maxit=60, ls_exact=1, **{"isreal": 0, **kw})
        Ug, _ = _compare(ctx, plan, op, b, (16, 16, 1), opts)
        if kw:
            assert np.abs(Ug.imag).max() == 0.0


@pytest.mark.parametrize("ls", [1, 0])
def test_ring_complex_solve(ctx, ls):
    geo = synth.ring_geometry(64, 64)
    op = O.Op.geometry(geo)
    x1, x2 = synth.disk_phantom(64, seed=1), synth.disk_phantom(64, seed=2)
    b = (synth.add_noise(op.gemv(x1), 0.005, 1) + 1j * synth.add_noise(op.gemv(x2), 0.005, 2))
    opts = tval3_amd.default_options(maxit=100, ls_exact=ls, isreal=0)
    _compare(ctx, ctx.plan_geometry(geo), op, b.astype(np.complex64), (64, 64, 1), opts,
             strict_traj=bool(ls))


@pytest.mark.parametrize("exact", [0, 1])
def test_hemisphere_complex_solve(ctx, exact):
    """3-D: the reference's Wz (y direction's ubar) and the fixed one."""
    grid = 16
    geo = synth.hemisphere_geometry(64, grid)
    op = O.Op.geometry(geo)
    x = synth.ellipsoid_phantom(grid, 6, seed=4)
    b = (synth.add_noise(op.gemv(x), 0.005, 4) * (0.6 + 0.8j)).astype(np.complex64)
    opts = tval3_amd.default_options(maxit=40, ls_exact=1, isreal=0, dt_exact=exact)
    _compare(ctx, ctx.plan_geometry(geo), op, b, (grid, grid, grid), opts)


def test_complex128_solves(ctx):
    A, b = _dense(15, t=np.complex128)
    opts = tval3_amd.default_options(maxit=60, ls_exact=1, isreal=0)
    _compare(ctx, ctx.plan_dense(A), O.Op.dense_c(A), b, (16, 16, 1), opts)
    geo = synth.ring_geometry(48, 48)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(48, seed=3)
    bg = (op.gemv(x) * (1 - 0.5j)).astype(np.complex128)
    _compare(ctx, ctx.plan_geometry(geo), op, bg, (48, 48, 1), opts)


def test_complex_batch_and_errors(ctx):
    A1, b1 = _dense(16)
    A2, b2 = _dense(17, m=120, n=256)
    p1, p2 = ctx.plan_dense(A1), ctx.plan_dense(A2)
    opts = tval3_amd.default_options(maxit=40, ls_exact=1, isreal=0)
    Us, infos = ctx.solve_batch([p1, p2, p1], [b1, b2, b1 * 0.5], [(16, 16, 1)] * 3, opts)
    for U, info, A, b in zip(Us, infos, (A1, A2, A1), (b1, b2, b1 * 0.5)):
        Uo, io, _ = O.solve_c(O.Op.dense_c(A), b, (16, 16, 1), opts)
        assert F.rel(U.cpu().numpy(), Uo) <= TOL_U and info.total_iters == io.total_iters
    with pytest.raises(tval3_amd.Tval3Error, match="tv3_solve_c64"):
        ctx.solve(ctx.plan_dense(A1.real.copy()), b1, (16, 16, 1), opts)
    with pytest.raises(tval3_amd.Tval3Error, match="own entry point"):
        ctx.solve(p1, b1.real.copy(), (16, 16, 1), opts)
    with pytest.raises(tval3_amd.Tval3Error, match="isotropic"):
        ctx.solve(p1, b1, (16, 16, 1), tval3_amd.default_options(maxit=5, tv_iso=1))


@pytest.mark.parametrize("t", [np.complex64, np.complex128])
def test_complex_graph_mode(ctx, t):
    """Steps replayed from a hipGraph give the host-driven complex result bit
    for bit, for a single solve (twice, a fresh capture each time), a batch
    and a ring geometry."""
    A, b = _dense(18, t=t)
    plan = ctx.plan_dense(A)
    kw = dict(maxit=60, ls_exact=1, isreal=0)
    U0, i0, _ = ctx.solve(plan, b, (16, 16, 1), tval3_amd.default_options(**kw))
    for rep in range(2):
        U1, i1, _ = ctx.solve(plan, b, (16, 16, 1), tval3_amd.default_options(graph=1, **kw))
        assert np.array_equal(U0.cpu().numpy(), U1.cpu().numpy()), rep
        assert (i1.total_iters, i1.outer_iters) == (i0.total_iters, i0.outer_iters)
    # a batch shares one operator kind: two dense instances of different size
    A2, b2 = _dense(19, m=120, n=256, t=t)
    plans, bs, dims = [plan, ctx.plan_dense(A2)], [b, b2], [(16, 16, 1), (16, 16, 1)]
    res = [ctx.solve_batch(plans, bs, dims, tval3_amd.default_options(graph=g, **kw))
           for g in (0, 1)]
    for k in range(2):
        assert np.array_equal(res[0][0][k].cpu().numpy(), res[1][0][k].cpu().numpy()), k
        assert res[0][1][k].total_iters == res[1][1][k].total_iters
    geo = synth.ring_geometry(32, 32)
    op = O.Op.geometry(geo)
    bg = (op.gemv(synth.disk_phantom(32, seed=7)) * (0.8 - 0.6j)).astype(t)
    gp = ctx.plan_geometry(geo)
    Ug = [ctx.solve(gp, bg, (32, 32, 1), tval3_amd.default_options(graph=g, **kw))[0].cpu().numpy()
          for g in (0, 1)]
    assert np.array_equal(Ug[0], Ug[1])
