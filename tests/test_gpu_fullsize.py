"""Parity at the sizes of the 3-D hemisphere and batched-dense workloads, and
the dense launch shapes the autotuner can pick.

- The 3-D hemisphere: 1024 Fibonacci transducers around a 128^3 volume
  (523,776 rays, about 32 M voxel visits per product), ellipsoid phantom with
  20 inclusions, 0.5 % noise, maxit 100: the GPU reconstruction and its
  per-iteration trajectory equal the oracle's bit for bit.
- The batched dense workload: 64 instances of uneven size (grids 64^2 ..
  256^2, m ~ U[20 %, 40 %] n, A ~ N(0, 1/m) stored bf16, 47.8 GiB of A), one
  batched solve with FP64 scalars, maxit 300.  The smallest instances equal
  the oracle bit for bit; every instance equals its own single-instance
  solve bit for bit (the two solves take their launch shapes from separate
  autotuner runs) and is finite.
- Every dense launch shape (TV3_DENSE_SHAPE=R,U,C: R rows per wavefront of
  A x, U rows loaded ahead, C columns per thread of the blocked A^T x) on a
  small uneven batch that runs all dense kernels, bit-identical to the
  oracle; and multi-layer rows that cross a layer boundary inside an R-row
  tile give the bits of R = 1.
"""
import numpy as np
import pytest

import tval3_amd
from tval3_amd import synth
import tv3_oracle as O

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _traj(tr):
    return [(t.iter, t.outer, t.backtracks, t.fallback, t.tau, t.alpha, t.f) for t in tr]


def test_hemisphere_c2_full_size(ctx):
    grid = 128
    geo = synth.hemisphere_geometry(1024, grid)
    plan = ctx.plan_geometry(geo)
    assert plan.m == 523776
    op = O.Op.geometry(geo)
    assert op.nnz == plan.nnz
    x = synth.ellipsoid_phantom(grid, 20, seed=0)
    y = ctx.gemv(plan, x).cpu().numpy()
    np.testing.assert_array_equal(y, op.gemv(x))
    b = synth.add_noise(y, 0.005, seed=0)
    dims = (grid, grid, grid)
    opts = tval3_amd.default_options(maxit=100)
    U, info, tr = ctx.solve(plan, b, dims, opts, Ut=x, trace_cap=100)
    Ug = U.cpu().numpy()
    Uo, info_o, tro = O.solve(op, b, dims, opts, Ut=x, trace_cap=100)
    print(f"hemisphere 128^3: iters {info.total_iters}/{info_o.total_iters} "
          f"outer {info.outer_iters}/{info_o.outer_iters} rel_error {info.rel_error:.4f}")
    assert np.all(np.isfinite(Ug))
    assert (info.total_iters, info.outer_iters) == (info_o.total_iters, info_o.outer_iters)
    assert _traj(tr) == _traj(tro)
    assert _bits_equal(Ug, Uo)


def _c3_batch(ctx):
    """The batched dense workload's instances, generated on the device (the
    same generator as the benchmark)."""
    import torch
    sizes = synth.batched_dense_sizes(64, seed=0)
    gen = torch.Generator(device=ctx.device)
    plans, bs, dims, mats = [], [], [], []
    for k, (g, m) in enumerate(sizes):
        n = g * g
        gen.manual_seed(1000 + k)
        Ab = torch.empty((m, n), dtype=torch.bfloat16, device=ctx.device)
        step = max(1, (1 << 28) // n)
        for r0 in range(0, m, step):
            r1 = min(m, r0 + step)
            Ab[r0:r1] = (torch.randn((r1 - r0, n), generator=gen, device=ctx.device)
                         * (1.0 / np.sqrt(m))).to(torch.bfloat16)
        x = torch.from_numpy(synth.ellipse_phantom(g, g, 8, seed=k)).to(ctx.device)
        y = torch.cat([Ab[r0:r0 + 4096].float() @ x for r0 in range(0, m, 4096)])
        b = torch.from_numpy(synth.add_noise(y.cpu().numpy(), 0.01, seed=k)).to(ctx.device)
        plans.append(ctx.plan_dense(Ab, bf16=True))
        bs.append(b)
        dims.append((g, g, 1))
        mats.append(Ab)
    return sizes, plans, bs, dims, mats


def test_dense_batch_c3_full_size(ctx):
    sizes, plans, bs, dims, mats = _c3_batch(ctx)
    tot = sum(g * g * m for g, m in sizes)
    print(f"C3 batch: 64 instances, {2 * tot / 2 ** 30:.1f} GiB of bf16 A")
    opts = tval3_amd.default_options(maxit=300, scal_double=1)
    Us, infos = ctx.solve_batch(plans, bs, dims, opts)
    Ub = [u.cpu().numpy() for u in Us]
    for k in range(64):
        assert np.all(np.isfinite(Ub[k])), k
        U1, info1, _ = ctx.solve(plans[k], bs[k], dims[k], opts)
        assert _bits_equal(U1.cpu().numpy(), Ub[k]), k
        assert info1.total_iters == infos[k].total_iters, k
    order = sorted(range(64), key=lambda k: sizes[k][0] ** 2 * sizes[k][1])[:3]
    for k in order:
        op = O.Op.dense(mats[k].float().cpu().numpy())
        Uo, info_o, _ = O.solve(op, bs[k].cpu().numpy(), dims[k], opts)
        assert info_o.total_iters == infos[k].total_iters, k
        assert _bits_equal(Ub[k], Uo), k


def _shape_batch(ctx, bf16):
    """Three instances that run every dense kernel: 128^2 with m = 4915 and
    1 MiB row blocks (308 blocks: the whole-row blocked A^T kernel), 40^2
    with m = 401 (few blocks: the row-slice pair) and 30^2 with m = 301 (rows
    of 900 entries: scalar loads); m is odd everywhere, so R-row tiles of
    A x end short."""
    import torch
    plans, ops, bs, dims = [], [], [], []
    for k, (g, m) in enumerate(((128, 4915), (40, 401), (30, 301))):
        n = g * g
        A = synth.gaussian_A(m, n, seed=70 + k)
        if bf16:
            At = torch.from_numpy(A).to(torch.bfloat16)
            A = At.to(torch.float32).numpy()
        else:
            At = A
        x = synth.ellipse_phantom(g, g, 8, seed=70 + k)
        b = synth.add_noise(A @ x, 0.01, seed=70 + k)
        plans.append(ctx.plan_dense(At, bf16=bf16))
        ops.append(O.Op.dense(A))
        bs.append(b)
        dims.append((g, g, 1))
    return plans, ops, bs, dims


@pytest.mark.parametrize("bf16", [True, False])
def test_every_dense_launch_shape(ctx, bf16, monkeypatch):
    plans, ops, bs, dims = _shape_batch(ctx, bf16)
    opts = tval3_amd.default_options(maxit=40, scal_double=1, submat_mb=1, ls_exact=1)
    want = [O.solve(op, b, d, opts)[0] for op, b, d in zip(ops, bs, dims)]
    for R in (1, 2, 4):
        for U in (1, 2, 4, 8):
            for Cc in (2, 4, 8):
                monkeypatch.setenv("TV3_DENSE_SHAPE", f"{R},{U},{Cc}")
                Us, _ = ctx.solve_batch(plans, bs, dims, opts)
                for k in range(3):
                    assert _bits_equal(Us[k].cpu().numpy(), want[k]), (R, U, Cc, k)


@pytest.mark.parametrize("bf16", [False, True])
def test_layered_rows_across_a_tile_boundary(ctx, bf16, monkeypatch):
    """3 layers of 16 x 16 sharing one A with m = 101 rows (n = 256, vector
    loads): the R-row tiles of the layered A x cross layer boundaries for
    R = 2 and 4, and those rows are summed in the order of R = 1."""
    import torch
    L, N, M = 3, 256, 101
    A = synth.gaussian_A(M, N, seed=88)
    At = torch.from_numpy(A).to(torch.bfloat16) if bf16 else A
    Ar = At.to(torch.float32).numpy() if bf16 else A
    x = np.concatenate([synth.ellipse_phantom(16, 16, 4, seed=88 + l).ravel() for l in range(L)])
    b = np.concatenate([synth.add_noise(Ar @ x[l * N:(l + 1) * N], 0.01, seed=88 + l)
                        for l in range(L)]).astype(np.float32)
    dims = (16, 16 * L, 1)
    plan = ctx.plan_dense(At, bf16=bf16)
    opts = tval3_amd.default_options(maxit=60, ls_exact=1)
    Uo, _, _ = O.solve(O.Op.dense(Ar), b, dims, opts)
    for R in (1, 2, 4):
        monkeypatch.setenv("TV3_DENSE_SHAPE", f"{R},4,8")
        U, _, _ = ctx.solve(plan, b, dims, opts)
        assert _bits_equal(U.cpu().numpy(), Uo), R


def _many_layers(L, q, p, m, seed, bf16):
    import torch
    N = q * p
    A = synth.gaussian_A(m, N, seed)
    At = torch.from_numpy(A).to(torch.bfloat16) if bf16 else A
    Ar = At.to(torch.float32).numpy() if bf16 else A
    x = np.concatenate([synth.ellipse_phantom(p, q, 4, seed + l).ravel() for l in range(L)])
    b = np.concatenate([synth.add_noise(Ar @ x[l * N:(l + 1) * N], 0.01, seed + l)
                        for l in range(L)]).astype(np.float32)
    return At, Ar, b, (q, p * L, 1)


# (layers, q, p, m, bf16): 64 layers with vector A rows and scalar m-vector
# loads (m % 4 = 1); 70 layers (a partial layer tile) of bf16 with vector
# loads on both sides; 20 layers of 15 x 13 (rows of 195 entries: scalar
# loads) with a row tile that ends short
MANY = [(64, 24, 24, 173, False), (70, 16, 16, 100, True), (20, 15, 13, 59, False)]


@pytest.mark.parametrize("L,q,p,m,bf16", MANY)
def test_layers_on_matrix_cores(ctx, L, q, p, m, bf16, monkeypatch):
    """The multi-layer products as FP64-MFMA GEMMs over the layers: the solve
    equals the oracle's bit for bit, with its trajectory, and equals the
    per-layer W = 1 kernels' solve (TV3_LAYER_MFMA=0) and the solves with
    one or 2 x 2 output tiles per workgroup in each product bit for bit."""
    At, Ar, b, dims = _many_layers(L, q, p, m, 90 + L, bf16)
    plan = ctx.plan_dense(At, bf16=bf16)
    opts = tval3_amd.default_options(maxit=60, ls_exact=1)
    U1, i1, t1 = ctx.solve(plan, b, dims, opts, trace_cap=60)
    Uo, io, to = O.solve(O.Op.dense(Ar), b, dims, opts, trace_cap=60)
    assert _bits_equal(U1.cpu().numpy(), Uo)
    assert _traj(t1) == _traj(to)
    monkeypatch.setenv("TV3_LAYER_MFMA", "0")
    U0, i0, _ = ctx.solve(plan, b, dims, opts)
    assert _bits_equal(U0.cpu().numpy(), U1.cpu().numpy())
    monkeypatch.delenv("TV3_LAYER_MFMA")
    # one 16 x 16 tile, 2 x 2 or (A X) 2 x 1 tiles per workgroup
    for blk in ("1,1", "2,2", "3,1", "2,1"):
        monkeypatch.setenv("TV3_MFMA_BLOCK", blk)
        Ub, _, _ = ctx.solve(plan, b, dims, opts)
        assert _bits_equal(Ub.cpu().numpy(), U1.cpu().numpy()), blk
    monkeypatch.delenv("TV3_MFMA_BLOCK")
    # maxit = 0 returns the scaled A^T b alone (the transposed GEMM)
    o0 = tval3_amd.default_options(maxit=0)
    Z1, _, _ = ctx.solve(plan, b, dims, o0)
    Zo, _, _ = O.solve(O.Op.dense(Ar), b, dims, o0)
    assert _bits_equal(Z1.cpu().numpy(), Zo)
    # a single-layer product on the same thread afterwards takes the W = 1
    # kernels again
    xl = synth.ellipse_phantom(p, q, 4, 5).ravel()
    np.testing.assert_array_equal(ctx.gemv(plan, xl).cpu().numpy(), O.Op.dense(Ar).gemv(xl))
