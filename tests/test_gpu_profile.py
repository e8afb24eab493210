"""The launch-timing report (tv3_ctx_profile levels 1 to 3 and
tv3_ctx_profile_slots), the counterpart of the reference's PROFILING table
(profile_info[slot], 3d/tval3_cpu_3d.cpp:1337-1348; 3d/mat_mult.h:19-44).

Timing must not change a result, the product slots must agree with
tv3_ctx_profile_read, and the slots carry the reference's numbers for the
passes that implement one reference function."""
import numpy as np
import pytest

import tval3_amd
from tval3_amd import synth
import tv3_oracle as O

pytestmark = pytest.mark.gpu


def _ring(grid=48, seed=3):
    geo = synth.ring_geometry(grid, grid)
    op = O.Op.geometry(geo)
    b = synth.add_noise(op.gemv(synth.disk_phantom(grid, seed=seed)), 0.005, seed)
    return geo, b


def _staged_transpose_bytes(geo, dims):
    """Operator bytes of the single-instance staged transpose: per tile of
    256 slots (16 x 16 squares of a slice, 8^3 bricks of a volume, voxels
    with 1 to 128 list entries) its ascending ray list when at most 8192
    rays cross it, the slots' lists as 2-byte positions in those lists
    (4-byte ray ids in tiles that stage none and in the longer lists, whose
    voxels also read their two list bounds), s_col, s_ptr and the tile
    pointers.  Restated here from the plan's definition, so the byte count
    also checks the tiles the library built."""
    op = O.Op.geometry(geo)
    ptr, lin = op.walk()
    m, nnz = len(ptr) - 1, len(lin)
    X, Y, Z = dims
    n = X * Y * Z
    b, bz, V, H = (8, 8, 256, 128) if Z > 1 else (16, 1, 256, 128)
    cnt = np.bincount(lin, minlength=n)
    v = np.arange(n)
    x, y, z = v % X, (v // X) % Y, v // (X * Y)
    key = np.lexsort((x, y, z, x // b, y // b, z // bz))
    nz = key[(cnt[key] > 0) & (cnt[key] <= H)]
    heavy = cnt > H
    slot = np.full(n, -1, np.int64)
    slot[nz] = np.arange(len(nz))
    ray = np.repeat(np.arange(m, dtype=np.int64), np.diff(ptr))
    sel = slot[lin] >= 0
    tile = slot[lin][sel] // V
    T = (len(nz) + V - 1) // V
    per_tile = np.bincount(np.unique(tile * m + ray[sel]) // m, minlength=T)
    ent = np.bincount(tile, minlength=T)
    staged = per_tile <= 8192
    s_rays = int(per_tile[staged].sum())
    dnnz = int(ent[~staged].sum()) + int(cnt[heavy].sum())
    return (2 * nnz + 2 * dnnz + 4 * n + 4 * (len(nz) + 1) + 8 * int(heavy.sum()) + 4 * (T + 1)
            + 4 * s_rays)


@pytest.mark.parametrize("ls", [1, 0])
def test_profile_slots_every_launch(ctx, ls):
    geo, b = _ring()
    plan = ctx.plan_geometry(geo)
    dims = (48, 48, 1)
    opts = tval3_amd.default_options(maxit=40, ls_exact=ls)
    U0, info0, _ = ctx.solve(plan, b, dims, opts)
    ref = U0.cpu().numpy()

    ctx.profile(2)
    U, info, _ = ctx.solve(plan, b, dims, opts)
    slots = ctx.profile_slots()
    read = ctx.profile_read()
    report = ctx.profile_report()
    ctx.profile(False)
    print(report)

    assert np.array_equal(U.cpu().numpy(), ref)
    assert info.total_iters == info0.total_iters
    d = {s: (name, runs, ms) for s, name, runs, ms in slots}
    assert [s for s, *_ in slots] == sorted(d)
    # scaleb, shrinkage, both products, descend, the setup / finish passes
    assert {1, 3, 10, 11, 30, 37} <= set(d), sorted(d)
    assert all(0 <= s < 40 and runs > 0 and ms > 0 and name for s, (name, runs, ms) in d.items())
    assert d[10][0].startswith("gemv CblasNoTrans") and d[11][0].startswith("gemv CblasConjTrans")
    # the product slots are the figures tv3_ctx_profile_read gives
    assert d[10][1] == read["A"][1] and d[11][1] == read["At"][1]
    assert d[10][2] == pytest.approx(read["A"][0], rel=1e-12)
    assert d[11][2] == pytest.approx(read["At"][0], rel=1e-12)
    # one shrinkage pass per inner iteration at least
    assert d[3][1] >= info.total_iters
    lines = report.splitlines()
    assert len(lines) == len(slots)
    assert lines[0].startswith(d[min(d)][0] + ":\t ") and "(runs: " in lines[0]


def test_profile_levels(ctx):
    geo, b = _ring(32, 4)
    plan = ctx.plan_geometry(geo)
    opts = tval3_amd.default_options(maxit=20, ls_exact=1)
    ctx.profile(1)
    ctx.solve(plan, b, (32, 32, 1), opts)
    s1 = ctx.profile_slots()
    ctx.profile(0)
    ctx.solve(plan, b, (32, 32, 1), opts)
    s0 = ctx.profile_slots()
    assert [s for s, *_ in s1] == [10, 11]
    assert s0 == []
    with pytest.raises(ValueError):
        ctx.profile(4)


def test_profile_level3_times_without_counting(ctx):
    """Level 3 times the same product launches as level 1 in a 64-lane batch
    but runs no counting kernel: only the launches the host counts (lanes
    gathered into single-instance kernels) add bytes and lanes."""
    geo, b = _ring(32, 4)
    plan = ctx.plan_geometry(geo)
    B = 70
    bs = [b * (1 + 0.01 * i) for i in range(B)]
    opts = tval3_amd.default_options(maxit=20, ls_exact=1)
    got = {}
    for level in (1, 3):
        ctx.profile(level)
        Us, _ = ctx.solve_batch([plan] * B, bs, [(32, 32, 1)] * B, opts)
        got[level] = (ctx.profile_read(), np.stack([u.cpu().numpy() for u in Us]))
        ctx.profile(0)
    (p1, U1), (p3, U3) = got[1], got[3]
    assert np.array_equal(U1, U3)
    for k in ("A", "At"):
        assert p3[k][1] == p1[k][1] > 0  # launches
        assert p1[k][2] > 0 and p1[k][3] > 0
    # the full 64-lane launches are counted on the device only
    assert p3["A"][2] < p1["A"][2] and p3["A"][3] < p1["A"][3]


def test_profile_batch_long_solve(ctx):
    """A batched solve long enough to fold the event pool in mid-solve: the
    per-slot counts match a level-1 run of the same solve."""
    geo, b = _ring(32, 5)
    plan = ctx.plan_geometry(geo)
    B = 8
    bs = [b * (1 + 0.01 * i) for i in range(B)]
    # tolerances no iterate meets: every instance runs all 800 iterations (about 8 launches each)
    opts = tval3_amd.default_options(maxit=800, ls_exact=1, tol=1e-30, tol_inn=1e-30)
    counts = {}
    for level in (1, 2):
        ctx.profile(level)
        Us, infos = ctx.solve_batch([plan] * B, bs, [(32, 32, 1)] * B, opts)
        counts[level] = {s: runs for s, _, runs, _ in ctx.profile_slots()}
        res = np.stack([u.cpu().numpy() for u in Us])
        ctx.profile(False)
        if level == 1:
            ref = res
        else:
            assert np.array_equal(res, ref)
    assert all(i.total_iters == 800 for i in infos)
    assert counts[2][10] == counts[1][10] and counts[2][11] == counts[1][11]
    assert sum(counts[2].values()) > 4096, counts[2]


@pytest.mark.parametrize("kind", ["f64", "c64", "c128"])
def test_profile_slots_variants(ctx, kind):
    """The double/double and complex solvers tag their launches with the
    same slots, and timing changes none of their results either."""
    geo, b = _ring(32, 6)
    plan = ctx.plan_geometry(geo)
    f64 = kind == "f64"
    bb = {"f64": b.astype(np.float64), "c64": (b * (0.6 + 0.8j)).astype(np.complex64),
          "c128": (b * (0.6 + 0.8j)).astype(np.complex128)}[kind]
    opts = tval3_amd.default_options(maxit=30, ls_exact=1, isreal=0 if kind != "f64" else 1)
    U0, i0, _ = ctx.solve(plan, bb, (32, 32, 1), opts, f64=f64)
    ctx.profile(2)
    U1, i1, _ = ctx.solve(plan, bb, (32, 32, 1), opts, f64=f64)
    d = {s: runs for s, _, runs, _ in ctx.profile_slots()}
    ctx.profile(False)
    assert np.array_equal(U0.cpu().numpy(), U1.cpu().numpy())
    assert i0.total_iters == i1.total_iters
    assert {1, 3, 10, 11, 30, 35, 37} <= set(d), sorted(d)
    assert d[3] >= i1.total_iters


@pytest.mark.parametrize("skip", [0, 1])
def test_profile_bytes_single_instance(ctx, skip):
    """A single-instance solve counts each product's compulsory bytes on the
    host (the A^T-skip launches on the device): every timed A launch of the
    projector carries its 32-byte ray descriptors and x and y (32 m + 4 n +
    4 m bytes), every A^T launch its staged voxel-major lists and vectors,
    and with skipping the A^T launches that skipped carry none."""
    geo, b = _ring(48, 7)
    plan = ctx.plan_geometry(geo)
    kw = dict(skipMulRatio=0.5, numInitIterations=3, maxRelativeChange=0.05) if skip else {}
    opts = tval3_amd.default_options(maxit=60, ls_exact=1, **kw)
    ctx.profile(1)
    ctx.solve(plan, b, (48, 48, 1), opts)
    r = ctx.profile_read()
    ctx.profile(False)
    m, n, nnz = plan.m, plan.n, plan.nnz
    per_a = 32 * m + 4 * n + 4 * m
    per_t = _staged_transpose_bytes(geo, (48, 48, 1)) + 4 * n + 4 * m
    assert r["A"][1] > 0 and r["A"][2] == r["A"][1] * per_a and r["A"][3] == r["A"][1]
    if skip:
        assert 0 < r["At"][2] <= r["At"][1] * per_t and r["At"][2] % per_t == 0
        assert r["At"][3] == r["At"][2] // per_t
        print("A^T launches", r["At"][1], "that computed", r["At"][2] // per_t)
    else:
        assert r["At"][2] == r["At"][1] * per_t and r["At"][3] == r["At"][1]


def test_profile_bytes_shared_batch(ctx):
    """70 instances sharing the ring geometry (two 64-lane groups): every A
    launch is charged the operator (ray descriptors and the 64-lane walk
    order, 36 m bytes) once per group that computed, and x and y once per
    computing instance, so the bytes minus the instances' vector bytes are
    a whole number of operator copies, at most two per launch.  A^T
    launches of lanes packed into single-instance vectors are charged the
    staged transpose's operator bytes instead of the 64-lane lists."""
    geo, b = _ring(32, 8)
    plan = ctx.plan_geometry(geo)
    B = 70
    bs = [b * (1 + 0.01 * i) for i in range(B)]
    opts = tval3_amd.default_options(maxit=30, ls_exact=1)
    ctx.profile(1)
    ctx.solve_batch([plan] * B, bs, [(32, 32, 1)] * B, opts)
    r = ctx.profile_read()
    ctx.profile(False)
    m, n, nnz = plan.m, plan.n, plan.nnz
    op1 = _staged_transpose_bytes(geo, (32, 32, 1))
    for key, op, alt in (("A", 36 * m, 32 * m), ("At", 4 * nnz + 8 * n + 4, op1)):
        ms, cnt, by, lanes = r[key]
        assert cnt > 0 and 0 < lanes <= cnt * B
        rest = by - lanes * 4 * (n + m)
        fits = [(i, j) for j in range(0, 2 * cnt + 1) for i in [(rest - j * alt) // op]
                if rest - j * alt >= 0 and (rest - j * alt) % op == 0 and cnt <= i + j <= 2 * cnt]
        assert fits, (key, rest / op, cnt)
