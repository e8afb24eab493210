"""Solver parity: full TVAL3 reconstructions on the GPU against the oracle.

Bar (from the workload definition): the FP32 reconstruction within 1e-4
relative L2 of the CPU result, with the per-iteration trajectory tracked.  The
kernels and the oracle share every elementwise rounding; sums differ only in
FP64 summation order, so the trajectories (iteration counts, backtracking
counts, fallbacks) are expected to coincide and the objective f to agree to a
few float ulps.  With ls_exact = 1 the backtracking runs the reference's rounds
and the trajectory must match exactly; the default quadratic line search
chooses the same step sizes except at near-ties of the Armijo test, so there
the reconstruction is held to the 1e-4 bar and the trajectory is reported.
"""
import numpy as np
import pytest

import tval3_amd
from tval3_amd import synth
import tv3_oracle as O

pytestmark = pytest.mark.gpu

TOL_U = 1e-4


def _rel(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _compare(ctx, plan, op, b, dims, opts, Ut=None, cap=400, strict_traj=True):
    U, info, tr = ctx.solve(plan, b, dims, opts, Ut=Ut, trace_cap=cap)
    Ug = U.cpu().numpy()
    Uo, info_o, tro = O.solve(op, b, dims, opts, Ut=Ut, trace_cap=cap)
    rel = _rel(Ug, Uo)
    same = [t.iter for t, s in zip(tr, tro)
            if (t.backtracks, t.fallback) != (s.backtracks, s.fallback)]
    msg = (f"rel={rel:.3e} iters gpu/oracle={info.total_iters}/{info_o.total_iters} "
           f"outer={info.outer_iters}/{info_o.outer_iters} first traj diff at {same[:3]} "
           f"backtracks gpu/oracle={sum(t.backtracks for t in tr)}/"
           f"{sum(t.backtracks for t in tro)}")
    print(msg)
    assert np.all(np.isfinite(Ug)), msg
    assert rel <= TOL_U, msg
    if strict_traj:
        assert info.total_iters == info_o.total_iters, msg
        assert info.outer_iters == info_o.outer_iters, msg
        assert not same, msg
        if tro:
            fo = np.array([s.f for s in tro]); fg = np.array([t.f for t in tr])
            assert np.allclose(fg, fo, rtol=1e-5, atol=1e-6 * np.abs(fo).max()), msg
    if Ut is not None:
        assert info.rel_error == pytest.approx(info_o.rel_error, rel=1e-3, abs=1e-6)
    return Ug, info, tr


def _dense_case(q, p, frac, seed):
    n = q * p
    m = int(round(frac * n))
    A = synth.gaussian_A(m, n, seed)
    x = synth.ellipse_phantom(p, q, 8, seed)
    b = synth.add_noise(A @ x, 0.01, seed)
    return A, x, b


@pytest.mark.parametrize("ls", [1, 0])
@pytest.mark.parametrize("sd", [0, 1])
@pytest.mark.parametrize("iso", [0, 1])
def test_dense_solve(ctx, sd, iso, ls):
    A, x, b = _dense_case(32, 32, 0.35, 11)
    opts = tval3_amd.default_options(maxit=120, scal_double=sd, tv_iso=iso, ls_exact=ls)
    _compare(ctx, ctx.plan_dense(A), O.Op.dense(A), b, (32, 32, 1), opts, Ut=x,
             strict_traj=bool(ls))


@pytest.mark.parametrize("ls", [1, 0])
def test_dense_c0_workload(ctx, ls):
    """64x64 image, m = 1229 (30 %), mu = 2^8, beta = 2^5, tol 1e-4, maxit 300."""
    A, x, b = _dense_case(64, 64, 0.3, 0)
    opts = tval3_amd.default_options(maxit=300, tv_iso=1, ls_exact=ls)
    _compare(ctx, ctx.plan_dense(A), O.Op.dense(A), b, (64, 64, 1), opts, Ut=x,
             strict_traj=bool(ls))


def test_dense_continuation_nonneg(ctx):
    A, x, b = _dense_case(24, 20, 0.5, 3)
    opts = tval3_amd.default_options(maxit=150, mu0=4.0, beta0=2.0, nonneg=1,
                                     tol_inn=1e-2, ls_exact=1)
    _compare(ctx, ctx.plan_dense(A), O.Op.dense(A), b, (24, 20, 1), opts, Ut=x)


def _layered(L, q, p_layer, frac, seed, sparse=False, n8=True):
    """L images of q x p_layer stacked along y and one A for all of them
    (the reference's 2-D multi-layer mode)."""
    N = q * p_layer
    M = int(round(frac * N))
    A = synth.gaussian_A(M, N, seed)
    if sparse:
        A[np.abs(A) < 0.8 * np.abs(A).mean()] = 0
    x = np.concatenate([synth.ellipse_phantom(p_layer, q, 4, seed + l).ravel() for l in range(L)])
    b = np.concatenate([synth.add_noise(A @ x[l * N:(l + 1) * N], 0.01, seed + l)
                        for l in range(L)]).astype(np.float32)
    return A, x, b, (q, L * p_layer, 1)


@pytest.mark.parametrize("ls", [1, 0])
@pytest.mark.parametrize("shape", [(3, 16, 16, 0.4), (2, 15, 13, 0.3026)])
def test_dense_layers(ctx, shape, ls, monkeypatch):
    """Multi-layer dense solve (layers = U pixels / A.n) against the oracle.
    15 x 13 layers have rows of 195 entries (scalar loads) and 59 rows, so
    with the untuned shape of two rows per wavefront one wavefront's rows
    cross the layer boundary."""
    monkeypatch.setenv("TV3_AUTOTUNE", "0")
    L, q, p, frac = shape
    A, x, b, dims = _layered(L, q, p, frac, 21)
    assert shape[0] != 2 or A.shape[0] == 59
    opts = tval3_amd.default_options(maxit=80, ls_exact=ls)
    _compare(ctx, ctx.plan_dense(A), O.Op.dense(A), b, dims, opts, Ut=x, strict_traj=bool(ls))


def test_dense_layers_bf16_and_sparse(ctx):
    import scipy.sparse as sp
    A, x, b, dims = _layered(2, 16, 12, 0.5, 31)
    A16 = A.astype(np.float32)
    Ab = ctx.torch.tensor(A16).to(ctx.torch.bfloat16).float().numpy()
    opts = tval3_amd.default_options(maxit=60, ls_exact=1, scal_double=1)
    _compare(ctx, ctx.plan_dense(A16, bf16=True), O.Op.dense(Ab), b, dims, opts, Ut=x)
    A, x, b, dims = _layered(3, 12, 10, 0.5, 41, sparse=True)
    S, T = sp.csr_matrix(A), sp.csc_matrix(A)
    args = (A.shape[0], A.shape[1], S.indptr, S.indices, S.data, T.indptr, T.indices, T.data)
    _compare(ctx, ctx.plan_sparse(*args), O.Op.sparse(*args), b, dims, opts, Ut=x)


def test_layers_batch_and_errors(ctx):
    """Multi-layer instances of different layer counts in one batch (W = 1
    groups); a dense batch cannot mix them with single-layer instances, and
    the projector takes no layers."""
    A, x, b, dims = _layered(3, 16, 16, 0.4, 51)
    A2, x2, b2, dims2 = _layered(2, 16, 16, 0.4, 52)
    plan = ctx.plan_dense(A)
    opts = tval3_amd.default_options(maxit=50, ls_exact=1)
    Us, infos = ctx.solve_batch([plan, plan], [b, b[: 2 * plan.m]], [dims, (16, 32, 1)], opts)
    for U, bb, d in ((Us[0], b, dims), (Us[1], b[: 2 * plan.m], (16, 32, 1))):
        Uo, _, _ = O.solve(O.Op.dense(A), bb, d, opts)
        assert _rel(U.cpu().numpy(), Uo) == 0.0
    b1 = b[: plan.m]
    with pytest.raises(tval3_amd.Tval3Error, match="mix"):
        ctx.solve_batch([plan, plan], [b, b1], [dims, (16, 16, 1)], opts)
    geo = synth.ring_geometry(16, 16)
    gplan = ctx.plan_geometry(geo)
    with pytest.raises(tval3_amd.Tval3Error, match="layers"):
        ctx.solve(gplan, np.zeros(2 * gplan.m, np.float32), (16, 32, 1), opts)


def _ring_case(K, grid, seed):
    geo = synth.ring_geometry(K, grid)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(grid, seed=seed)
    b = synth.add_noise(op.gemv(x), 0.005, seed)
    return geo, op, x, b


@pytest.mark.parametrize("ls", [1, 0])
@pytest.mark.parametrize("sd", [0, 1])
def test_ring_geometry_solve(ctx, sd, ls):
    geo, op, x, b = _ring_case(64, 64, 1)
    opts = tval3_amd.default_options(maxit=100, scal_double=sd, ls_exact=ls)
    _compare(ctx, ctx.plan_geometry(geo), op, b, (64, 64, 1), opts, Ut=x,
             strict_traj=bool(ls))


def test_multi_mp_geometry(ctx):
    """Three measurement positions (numMPs > 1) with skipped pairs, a shuffled
    output order and rows no pair maps to: products, one solve and a 70-lane
    batch (two 64-lane groups, tiled walk orders over all positions) against
    the oracle."""
    geo = synth.multi_mp_ring_geometry(32, 40, 3, skip_every=5, extra_rows=9, seed=3)
    op = O.Op.geometry(geo)
    plan = ctx.plan_geometry(geo)
    x = synth.disk_phantom(40, seed=4)
    y = ctx.gemv(plan, x).cpu().numpy()
    np.testing.assert_array_equal(y, op.gemv(x))
    yb = synth.add_noise(y, 0.01, 4)
    np.testing.assert_array_equal(ctx.gemv(plan, yb, trans=True).cpu().numpy(), op.gemv(yb, trans=True))
    opts = tval3_amd.default_options(maxit=60, ls_exact=1)
    _compare(ctx, plan, op, yb, (40, 40, 1), opts, Ut=x)
    B = 70
    bs = [synth.add_noise(y, 0.01, 10 + i) for i in range(B)]
    Us, infos = ctx.solve_batch([plan] * B, bs, [(40, 40, 1)] * B, opts)
    for i in (0, 33, 64, 69):
        Uo, info_o, _ = O.solve(op, bs[i], (40, 40, 1), opts)
        assert _rel(Us[i].cpu().numpy(), Uo) == 0.0, i
        assert infos[i].total_iters == info_o.total_iters


@pytest.mark.parametrize("ls", [1, 0])
def test_ring256_workload(ctx, ls):
    """The benchmark workload itself (256 transducers, 256x256, maxit 200)."""
    geo, op, x, b = _ring_case(256, 256, 0)
    opts = tval3_amd.default_options(maxit=200, ls_exact=ls)
    _compare(ctx, ctx.plan_geometry(geo), op, b, (256, 256, 1), opts, Ut=x,
             strict_traj=bool(ls))


def test_ring_sparse_solve(ctx):
    geo, op, x, b = _ring_case(64, 64, 2)
    ptr, lin = op.walk()
    csr = synth.csr_from_walk(ptr, lin, geo.n)
    sop = O.Op.sparse(geo.m, geo.n, *csr)
    opts = tval3_amd.default_options(maxit=100, ls_exact=1)
    _compare(ctx, ctx.plan_sparse(geo.m, geo.n, *csr), sop, b, (64, 64, 1), opts, Ut=x)


def test_ring_skip_mul(ctx):
    geo, op, x, b = _ring_case(64, 64, 3)
    opts = tval3_amd.default_options(maxit=80, skipMulRatio=0.5,
                                     numInitIterations=3, maxRelativeChange=0.05,
                                     ls_exact=1)
    _compare(ctx, ctx.plan_geometry(geo), op, b, (64, 64, 1), opts, Ut=x)


@pytest.mark.parametrize("exact", [0, 1])
def test_hemisphere_3d_solve(ctx, exact):
    grid = 20
    geo = synth.hemisphere_geometry(96, grid)
    op = O.Op.geometry(geo)
    x = synth.ellipsoid_phantom(grid, 6, seed=4)
    b = synth.add_noise(op.gemv(x), 0.005, 4)
    opts = tval3_amd.default_options(maxit=40, dt_exact=exact, ls_exact=1)
    _compare(ctx, ctx.plan_geometry(geo), op, b, (grid, grid, grid), opts, Ut=x)


def _shared_ring_batch(ctx, B, opts, seed0=100):
    geo, op, x, _ = _ring_case(48, 48, 5)
    plan = ctx.plan_geometry(geo)
    bs = [synth.add_noise(op.gemv(x), 0.005, seed0 + i) for i in range(B)]
    Us, infos = ctx.solve_batch([plan] * B, bs, [(48, 48, 1)] * B, opts, Uts=[x] * B)
    return op, x, bs, np.stack([u.cpu().numpy() for u in Us]), infos


@pytest.mark.parametrize("B,ls", [(70, 1), (130, 1), (130, 0)])
def test_batch_shared_ring(ctx, B, ls, monkeypatch):
    """Instances sharing one ring geometry in 64-lane groups (70: two groups,
    the last padded; 130: three).  Most steps have instances in the 5-strike
    fallback or restarting their inner loop; their masked A / A^T products run
    on the full groups, on compact 64-lane groups or on W = 1 lanes, and all
    three give the same bits.  A sample of instances matches the oracle."""
    opts = tval3_amd.default_options(maxit=60, ls_exact=ls)
    res = {}
    for mode, lane_max, c64 in [("full", "0", "0"), ("compact64", "0", "1"), ("w1", "64", "0")]:
        monkeypatch.setenv("TV3_LANE_MAX", lane_max)
        monkeypatch.setenv("TV3_COMPACT64", c64)
        op, x, bs, U, infos = _shared_ring_batch(ctx, B, opts)
        res[mode] = (U, [i.total_iters for i in infos])
    for mode in ("compact64", "w1"):
        assert np.array_equal(res[mode][0], res["full"][0]), mode
        assert res[mode][1] == res["full"][1], mode
    U, iters = res["compact64"]
    for i in sorted({0, 1, 15, 16, 63, 64, B - 1}):
        Uo, info_o, _ = O.solve(op, bs[i], (48, 48, 1), opts, Ut=x)
        rel = _rel(U[i], Uo)
        assert rel <= TOL_U, (i, rel, iters[i], info_o.total_iters)
        if ls:
            assert iters[i] == info_o.total_iters


def test_batch_shared_ring_16_lane_groups(ctx, monkeypatch):
    """TV3_GROUP_WIDTH=16: the same instances in 16-lane groups give the bits
    of the 64-lane groups."""
    opts = tval3_amd.default_options(maxit=60, ls_exact=1)
    _, _, _, U64, _ = _shared_ring_batch(ctx, 70, opts)
    monkeypatch.setenv("TV3_GROUP_WIDTH", "16")
    _, _, _, U16, _ = _shared_ring_batch(ctx, 70, opts)
    assert np.array_equal(U16, U64)


def test_batch_uneven_dense(ctx):
    sizes = [(16, 16, 0.3), (24, 20, 0.4), (32, 32, 0.25)]
    plans, bs, dims, cases = [], [], [], []
    for k, (q, p, fr) in enumerate(sizes):
        A, x, b = _dense_case(q, p, fr, 20 + k)
        plans.append(ctx.plan_dense(A)); bs.append(b); dims.append((q, p, 1))
        cases.append((A, x, b))
    opts = tval3_amd.default_options(maxit=80)
    Us, infos = ctx.solve_batch(plans, bs, dims, opts)
    for k, (A, x, b) in enumerate(cases):
        Uo, info_o, _ = O.solve(O.Op.dense(A), b, dims[k], opts)
        assert _rel(Us[k].cpu().numpy(), Uo) <= TOL_U
        assert infos[k].total_iters == info_o.total_iters


def test_maxit_zero_returns_scaled_Atb(ctx):
    A, x, b = _dense_case(16, 16, 0.5, 9)
    opts = tval3_amd.default_options(maxit=0)
    _compare(ctx, ctx.plan_dense(A), O.Op.dense(A), b, (16, 16, 1), opts)


def test_argument_errors(ctx):
    A, x, b = _dense_case(16, 16, 0.5, 9)
    plan = ctx.plan_dense(A)
    with pytest.raises(tval3_amd.Tval3Error, match="length of U"):
        ctx.solve(plan, b, (15, 16, 1))
    geo = synth.ring_geometry(16, 16)
    geo.xe[0] = 99
    with pytest.raises(tval3_amd.Tval3Error, match="outside"):
        ctx.plan_geometry(geo)


def test_batch_bf16_dense_uneven(ctx):
    """Uneven-size batch with A stored in bf16 (the batched dense workload at
    small sizes).  The kernels widen bf16 exactly, so the oracle run on the
    bf16-rounded A (as float32) is the parity reference; the FP64 solve of the
    same matrices gives the reported error."""
    import torch
    sizes = [(24, 24, 0.3), (32, 32, 0.25), (20, 28, 0.4), (40, 40, 0.2)]
    plans, bs, dims, cases = [], [], [], []
    for k, (q, p, fr) in enumerate(sizes):
        A, x, b = _dense_case(q, p, fr, 40 + k)
        Ab = torch.from_numpy(A).to(torch.bfloat16)
        A_r = Ab.to(torch.float32).numpy()
        b = synth.add_noise(A_r @ x, 0.01, 40 + k)
        plans.append(ctx.plan_dense(Ab, bf16=True)); bs.append(b); dims.append((q, p, 1))
        cases.append((A_r, x, b))
    opts = tval3_amd.default_options(maxit=100, ls_exact=1)
    Us, infos = ctx.solve_batch(plans, bs, dims, opts, Uts=[c[1] for c in cases])
    for k, (A_r, x, b) in enumerate(cases):
        op = O.Op.dense(A_r)
        Uo, info_o, _ = O.solve(op, b, dims[k], opts, Ut=x)
        Ug = Us[k].cpu().numpy()
        assert _rel(Ug, Uo) <= TOL_U, k
        assert infos[k].total_iters == info_o.total_iters
        U64, _, _ = O.solve_f64(op, b, dims[k], opts, Ut=x)
        print(f"instance {k}: rel L2 vs FP64 solve {_rel(Ug, U64):.2e}")
        assert infos[k].rel_error == pytest.approx(info_o.rel_error, rel=1e-3)


# ---- graph mode: the same launches, captured once per two steps --------------
def _same(a, b):
    return np.array_equal(a.cpu().numpy(), b.cpu().numpy())


@pytest.mark.parametrize("ls", [0, 1])
def test_graph_mode_plain_ring(ctx, ls):
    """Uniform masked steps replayed from a hipGraph give the host-driven
    result bit for bit (every replay feeds the next, so the final iterate
    checks them all) and match the oracle's trajectory."""
    geo, op, x, b = _ring_case(48, 48, 5)
    plan = ctx.plan_geometry(geo)
    o_host = tval3_amd.default_options(maxit=120, ls_exact=ls)
    o_graph = tval3_amd.default_options(maxit=120, ls_exact=ls, graph=1)
    U0, i0, t0 = ctx.solve(plan, b, (48, 48, 1), o_host, Ut=x, trace_cap=120)
    for rep in range(2):  # two solves in a row: a fresh capture each time
        U1, i1, t1 = ctx.solve(plan, b, (48, 48, 1), o_graph, Ut=x, trace_cap=120)
        assert _same(U0, U1), rep
        assert (i1.total_iters, i1.outer_iters) == (i0.total_iters, i0.outer_iters)
        assert [(t.backtracks, t.fallback, t.tau) for t in t1] == \
               [(t.backtracks, t.fallback, t.tau) for t in t0]
    _compare(ctx, plan, op, b, (48, 48, 1), o_graph, Ut=x, strict_traj=(ls == 1))


def test_graph_mode_shared_batch(ctx):
    geo, op, x, _ = _ring_case(48, 48, 5)
    plan = ctx.plan_geometry(geo)
    B = 70
    bs = [synth.add_noise(op.gemv(x), 0.005, 300 + i) for i in range(B)]
    res = []
    for g in (0, 1):
        opts = tval3_amd.default_options(maxit=60, ls_exact=1, graph=g)
        res.append(ctx.solve_batch([plan] * B, bs, [(48, 48, 1)] * B, opts))
    (Ua, ia), (Ub, ib) = res
    for k in range(B):
        assert _same(Ua[k], Ub[k]), k
        assert ia[k].total_iters == ib[k].total_iters
    Uo, info_o, _ = O.solve(op, bs[5], (48, 48, 1), tval3_amd.default_options(maxit=60, ls_exact=1))
    assert _rel(Ub[5].cpu().numpy(), Uo) <= TOL_U


def test_graph_mode_dense_bf16_batch_and_continuation(ctx):
    import torch
    sizes = [(24, 24, 0.3), (20, 28, 0.4), (32, 32, 0.25)]
    plans, bs, dims = [], [], []
    for k, (q, p, fr) in enumerate(sizes):
        A, x, b = _dense_case(q, p, fr, 60 + k)
        plans.append(ctx.plan_dense(torch.from_numpy(A).to(torch.bfloat16), bf16=True))
        bs.append(b)
        dims.append((q, p, 1))
    kw = dict(maxit=150, mu0=16.0, beta0=4.0, nonneg=1, scal_double=1)
    Ua, ia = ctx.solve_batch(plans, bs, dims, tval3_amd.default_options(**kw))
    Ub, ib = ctx.solve_batch(plans, bs, dims, tval3_amd.default_options(graph=1, **kw))
    for k in range(len(sizes)):
        assert _same(Ua[k], Ub[k]), k
        assert (ia[k].total_iters, ia[k].outer_iters) == (ib[k].total_iters, ib[k].outer_iters)
