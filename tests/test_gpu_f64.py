"""The double/double path (tv3_solve_f64, tv3_gemv_f64): FP64 vectors and
scalars, the reference's mat<double> / sparse_mat<double> / geometry
overloads with T_comp = double.

Bar: products within 1e-13 relative of an FP64 evaluation (numpy's for
double-valued matrices, the oracle's for the projector).  For whole solves
both sides sum in FP64 but in different orders, and the TVAL3 iteration
amplifies rounding differences about tenfold every five inner iterations
(the oracle's own result moves by up to 3e-6 on the 64x64 ring when only the
chunk length of its FP64 dot products changes).  So a solve is held to:
  * the same trajectory (iteration, outer and backtracking counts);
  * the first five iterations' f and tau within 1e-12 relative (there the
    two runs differ by rounding only, so a wrong term shows at once);
  * the final U within 10x the oracle's own spread under other reduction
    orders and rounding-level perturbations of b (floor 1e-12; f64_bar).
"""
import numpy as np
import pytest

import tval3_amd
from tval3_amd import synth
import tv3_oracle as O
from f64_bar import rel as _rel, bound as _bound

pytestmark = pytest.mark.gpu

TOL_OP = 1e-13
TOL_EARLY = 1e-12


def _compare(ctx, plan, op, b, dims, opts, Ut=None, cap=400):
    U, info, tr = ctx.solve(plan, b, dims, opts, Ut=Ut, trace_cap=cap, f64=True)
    assert U.dtype == ctx.torch.float64
    Ug = U.cpu().numpy()
    Uo, info_o, tro = O.solve_f64(op, b, dims, opts, Ut=Ut, trace_cap=cap)
    rel = _rel(Ug, Uo)
    bar = _bound(op, b, dims, opts, Uo)
    diff = [t.iter for t, s in zip(tr, tro)
            if (t.backtracks, t.fallback) != (s.backtracks, s.fallback)]
    early = max([abs(t.f - s.f) / max(abs(s.f), 1e-300) for t, s in zip(tr[:5], tro[:5])] +
                [abs(t.tau - s.tau) / max(abs(s.tau), 1e-300) for t, s in zip(tr[:5], tro[:5])])
    msg = (f"rel={rel:.3e} bar={bar:.3e} early={early:.2e} "
           f"iters gpu/oracle={info.total_iters}/{info_o.total_iters} "
           f"outer={info.outer_iters}/{info_o.outer_iters} first traj diff at {diff[:3]}")
    print(msg)
    assert np.all(np.isfinite(Ug)), msg
    assert info.total_iters == info_o.total_iters, msg
    assert info.outer_iters == info_o.outer_iters, msg
    assert not diff, msg
    assert len(tr) >= min(5, info_o.total_iters), msg
    assert early <= TOL_EARLY, msg
    assert rel <= bar, msg
    return Ug, info


def _dense_case(q, p, frac, seed):
    n = q * p
    m = int(round(frac * n))
    A = synth.gaussian_A(m, n, seed)
    x = synth.ellipse_phantom(p, q, 8, seed)
    b = synth.add_noise(A @ x, 0.01, seed).astype(np.float64)
    return A, x, b


def test_gemv_f64_double_valued(ctx):
    """Dense and sparse matrices with entries that are not floats: y = A x
    and A^T y against numpy's FP64 products (the oracle holds float entries
    only, so these two are pinned by numpy)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(5)
    for m, n in ((37, 53), (300, 1000), (1, 7)):
        A = rng.standard_normal((m, n))
        x = rng.standard_normal(n)
        y = rng.standard_normal(m)
        plan = ctx.plan_dense(A, f64=True)
        assert _rel(ctx.gemv(plan, x, f64=True).cpu().numpy(), A @ x) <= TOL_OP
        assert _rel(ctx.gemv(plan, y, trans=True, f64=True).cpu().numpy(), A.T @ y) <= TOL_OP
        As = A * (rng.random((m, n)) < 0.2)
        S, T = sp.csr_matrix(As), sp.csc_matrix(As)
        splan = ctx.plan_sparse(m, n, S.indptr, S.indices, S.data, T.indptr, T.indices,
                                T.data, f64=True)
        assert _rel(ctx.gemv(splan, x, f64=True).cpu().numpy(), As @ x) <= TOL_OP
        assert _rel(ctx.gemv(splan, y, trans=True, f64=True).cpu().numpy(), As.T @ y) <= TOL_OP


def test_gemv_f64_geometry(ctx):
    """The projector on double vectors against the oracle's gemv<double>
    (products x * (double)scale, FP64 sums): 2-D ring, numMPs > 1 with
    skipped pairs, 3-D hemisphere."""
    rng = np.random.default_rng(6)
    for geo in (synth.ring_geometry(64, 64),
                synth.multi_mp_ring_geometry(32, 40, 3, skip_every=5, extra_rows=9, seed=3),
                synth.hemisphere_geometry(64, 16)):
        op = O.Op.geometry(geo)
        plan = ctx.plan_geometry(geo)
        x = rng.standard_normal(geo.n)
        y = rng.standard_normal(geo.m)
        assert _rel(ctx.gemv(plan, x, f64=True).cpu().numpy(), op.gemv_f64(x)) <= TOL_OP
        assert _rel(ctx.gemv(plan, y, trans=True, f64=True).cpu().numpy(),
                    op.gemv_f64(y, trans=True)) <= TOL_OP


@pytest.mark.parametrize("ls", [1, 0])
@pytest.mark.parametrize("iso", [0, 1])
def test_dense_solve_f64(ctx, iso, ls):
    A, x, b = _dense_case(32, 32, 0.35, 11)
    opts = tval3_amd.default_options(maxit=120, tv_iso=iso, ls_exact=ls)
    _compare(ctx, ctx.plan_dense(A, f64=True), O.Op.dense(A), b, (32, 32, 1), opts, Ut=x)


def test_dense_continuation_nonneg_f64(ctx):
    A, x, b = _dense_case(24, 20, 0.5, 3)
    opts = tval3_amd.default_options(maxit=150, mu0=4.0, beta0=2.0, nonneg=1,
                                     tol_inn=1e-2, ls_exact=1)
    _compare(ctx, ctx.plan_dense(A, f64=True), O.Op.dense(A), b, (24, 20, 1), opts, Ut=x)


def test_dense_layers_f64(ctx):
    """2-D multi-layer mode with double vectors."""
    L, q, p = 3, 16, 12
    N = q * p
    A = synth.gaussian_A(int(0.4 * N), N, 21)
    x = np.concatenate([synth.ellipse_phantom(p, q, 4, 21 + l).ravel() for l in range(L)])
    b = np.concatenate([synth.add_noise(A @ x[l * N:(l + 1) * N], 0.01, 21 + l)
                        for l in range(L)]).astype(np.float64)
    opts = tval3_amd.default_options(maxit=60, ls_exact=1)
    _compare(ctx, ctx.plan_dense(A, f64=True), O.Op.dense(A), b, (q, L * p, 1), opts, Ut=x)


@pytest.mark.parametrize("ls", [1, 0])
def test_ring_geometry_solve_f64(ctx, ls):
    geo = synth.ring_geometry(64, 64)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(64, seed=1)
    b = synth.add_noise(op.gemv(x), 0.005, 1).astype(np.float64)
    opts = tval3_amd.default_options(maxit=100, ls_exact=ls)
    _compare(ctx, ctx.plan_geometry(geo), op, b, (64, 64, 1), opts, Ut=x)


def test_ring_sparse_and_skip_f64(ctx):
    """sparse_mat<double> from the ring's walk, and A^T skipping (the skip
    policy) on the projector."""
    geo = synth.ring_geometry(48, 48)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(48, seed=2)
    b = synth.add_noise(op.gemv(x), 0.005, 2).astype(np.float64)
    ptr, lin = op.walk()
    csr = synth.csr_from_walk(ptr, lin, geo.n)
    sop = O.Op.sparse(geo.m, geo.n, *csr)
    opts = tval3_amd.default_options(maxit=80, ls_exact=1)
    _compare(ctx, ctx.plan_sparse(geo.m, geo.n, *csr, f64=True), sop, b, (48, 48, 1), opts, Ut=x)
    opts = tval3_amd.default_options(maxit=80, skipMulRatio=0.5, numInitIterations=3,
                                     maxRelativeChange=0.05, ls_exact=1)
    _compare(ctx, ctx.plan_geometry(geo), op, b, (48, 48, 1), opts, Ut=x)


def test_hemisphere_3d_solve_f64(ctx):
    grid = 16
    geo = synth.hemisphere_geometry(64, grid)
    op = O.Op.geometry(geo)
    x = synth.ellipsoid_phantom(grid, 6, seed=4)
    b = synth.add_noise(op.gemv(x), 0.005, 4).astype(np.float64)
    opts = tval3_amd.default_options(maxit=40, ls_exact=1)
    _compare(ctx, ctx.plan_geometry(geo), op, b, (grid, grid, grid), opts, Ut=x)


def test_batch_f64(ctx):
    """Three instances of different operators and sizes in one batch (one
    instance per group) against separate oracle solves."""
    A1, x1, b1 = _dense_case(16, 16, 0.4, 7)
    A2, x2, b2 = _dense_case(20, 12, 0.5, 8)
    p1, p2 = ctx.plan_dense(A1, f64=True), ctx.plan_dense(A2, f64=True)
    opts = tval3_amd.default_options(maxit=60, ls_exact=1)
    dims = [(16, 16, 1), (20, 12, 1), (16, 16, 1)]
    bs = [b1, b2, b1 * 0.5]
    Us, infos = ctx.solve_batch([p1, p2, p1], bs, dims, opts, f64=True)
    for U, info, A, b, d in zip(Us, infos, (A1, A2, A1), bs, dims):
        op = O.Op.dense(A)
        Uo, info_o, _ = O.solve_f64(op, b, d, opts)
        assert _rel(U.cpu().numpy(), Uo) <= _bound(op, b, d, opts, Uo)
        assert info.total_iters == info_o.total_iters


def test_f64_plan_errors(ctx):
    """Float entry points reject double-valued plans and the double ones
    float-valued plans; the projector serves both."""
    A, x, b = _dense_case(16, 16, 0.4, 9)
    opts = tval3_amd.default_options(maxit=5)
    p32, p64 = ctx.plan_dense(A), ctx.plan_dense(A, f64=True)
    with pytest.raises(tval3_amd.Tval3Error, match="tv3_solve_f64"):
        ctx.solve(p64, b, (16, 16, 1), opts)
    with pytest.raises(tval3_amd.Tval3Error, match="tv3_solve_f64"):
        ctx.solve(p32, b, (16, 16, 1), opts, f64=True)
    with pytest.raises(tval3_amd.Tval3Error, match="tv3_gemv_f64"):
        ctx.gemv(p64, x.ravel())
    with pytest.raises(tval3_amd.Tval3Error, match="tv3_gemv_f64"):
        ctx.gemv(p32, x.ravel(), f64=True)


@pytest.mark.parametrize("ls", [1, 0])
def test_graph_mode_f64(ctx, ls):
    """Steps replayed from a hipGraph give the host-driven FP64 result bit
    for bit: one ring solve twice in a row (a fresh capture each time) and a
    batch of two operators."""
    geo = synth.ring_geometry(48, 48)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(48, seed=6)
    b = synth.add_noise(op.gemv(x), 0.005, 6).astype(np.float64)
    plan = ctx.plan_geometry(geo)
    kw = dict(maxit=100, ls_exact=ls)
    U0, i0, t0 = ctx.solve(plan, b, (48, 48, 1), tval3_amd.default_options(**kw), Ut=x,
                           trace_cap=100, f64=True)
    for rep in range(2):
        U1, i1, t1 = ctx.solve(plan, b, (48, 48, 1), tval3_amd.default_options(graph=1, **kw),
                               Ut=x, trace_cap=100, f64=True)
        assert np.array_equal(U0.cpu().numpy(), U1.cpu().numpy()), rep
        assert (i1.total_iters, i1.outer_iters) == (i0.total_iters, i0.outer_iters)
        assert [(t.backtracks, t.fallback, t.tau) for t in t1] == \
               [(t.backtracks, t.fallback, t.tau) for t in t0]
    # a batch shares one operator kind: two dense instances of different size
    A1, _, b1 = _dense_case(16, 16, 0.4, 11)
    A2, _, b2 = _dense_case(20, 12, 0.5, 12)
    plans = [ctx.plan_dense(A1, f64=True), ctx.plan_dense(A2, f64=True)]
    bs, dims = [b1, b2], [(16, 16, 1), (20, 12, 1)]
    res = [ctx.solve_batch(plans, bs, dims, tval3_amd.default_options(graph=g, **kw), f64=True)
           for g in (0, 1)]
    for k in range(2):
        assert np.array_equal(res[0][0][k].cpu().numpy(), res[1][0][k].cpu().numpy()), k
        assert res[0][1][k].total_iters == res[1][1][k].total_iters


def test_bf16_A_with_f64_vectors(ctx):
    """A bf16-stored A with FP64 vectors and scalars: the FP64 solve of the
    bf16-rounded matrix.  An uneven batch (vector-aligned rows and one 15 x 13
    instance with rows of 195 entries) agrees with the same solves on FP64
    copies of the rounded matrices to FP64 rounding (their products sum in a
    different order), and each solve meets the FP64 bar against the
    oracle's double/double solve of the rounded matrix."""
    import torch
    sizes = [(24, 24, 0.3), (32, 32, 0.25), (20, 28, 0.4), (15, 13, 0.5)]
    p16, p64, bs, dims, mats = [], [], [], [], []
    for k, (q, p, fr) in enumerate(sizes):
        A, x, _ = _dense_case(q, p, fr, 70 + k)
        Ab = torch.from_numpy(A).to(torch.bfloat16)
        Ar = Ab.to(torch.float64).numpy()
        b = synth.add_noise((Ar @ x.ravel()).astype(np.float32), 0.01, 70 + k).astype(np.float64)
        p16.append(ctx.plan_dense(Ab, bf16=True))
        p64.append(ctx.plan_dense(Ar, f64=True))
        bs.append(b)
        dims.append((q, p, 1))
        mats.append(Ar)
    opts = tval3_amd.default_options(maxit=80, ls_exact=1)
    U16, i16 = ctx.solve_batch(p16, bs, dims, opts, f64=True)
    U64, i64 = ctx.solve_batch(p64, bs, dims, opts, f64=True)
    for k in range(len(sizes)):
        u16, u64 = U16[k].cpu().numpy(), U64[k].cpu().numpy()
        op = O.Op.dense(mats[k].astype(np.float32))
        Uo, info_o, _ = O.solve_f64(op, bs[k], dims[k], opts)
        bar = _bound(op, bs[k], dims[k], opts, Uo)
        assert i16[k].total_iters == i64[k].total_iters == info_o.total_iters, k
        assert _rel(u16, Uo) <= bar and _rel(u64, Uo) <= bar, (k, _rel(u16, Uo), bar)
        print(f"instance {k}: bf16-A FP64 vs f64-A FP64 {_rel(u16, u64):.2e} (bar {bar:.2e})")
    # the product itself
    xk = np.random.default_rng(1).standard_normal(mats[0].shape[1])
    y16 = ctx.gemv(p16[0], xk, f64=True).cpu().numpy()
    np.testing.assert_allclose(y16, mats[0] @ xk, rtol=1e-13, atol=1e-13 * np.abs(mats[0]).sum(1).max())
    yt = ctx.gemv(p16[3], bs[3], trans=True, f64=True).cpu().numpy()
    np.testing.assert_allclose(yt, mats[3].T @ bs[3], rtol=1e-13, atol=1e-13 * np.abs(mats[3]).sum(0).max())
