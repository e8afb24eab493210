"""tv3_gemv_noalloc, the gemv entry point that allocates nothing, against
tv3_gemv through the C-ABI (ctypes, device pointers from torch tensors).

Both run the same kernels on the same inputs, so every output must be
bit-identical to tv3_gemv's.  Each case makes two calls in a row on one
context, the second with a smaller operator, so the second call runs on the
fixed device scratch the first one left behind (the dense A^T x partial sums
and the group descriptor are rewritten by every call).
"""
import ctypes as C

import numpy as np
import pytest

import tval3_amd
from tval3_amd import synth

pytestmark = pytest.mark.gpu


def _call(ctx, name, plan, trans, x, y):
    L = tval3_amd.lib()
    ctx._enter()
    rc = getattr(L, name)(ctx.h, plan.h, int(trans), C.c_void_p(x.data_ptr()),
                          C.c_void_p(y.data_ptr()))
    ctx._exit()
    return rc


def _pair(ctx, plan, trans, seed):
    t = ctx.torch
    rng = np.random.default_rng(seed)
    nin, nout = (plan.m, plan.n) if trans else (plan.n, plan.m)
    x = ctx.tensor(rng.standard_normal(nin).astype(np.float32))
    y_ref = t.full((nout,), float("nan"), dtype=t.float32, device=ctx.device)
    y_na = t.full((nout,), float("nan"), dtype=t.float32, device=ctx.device)
    assert _call(ctx, "tv3_gemv", plan, trans, x, y_ref) == 0
    rc = _call(ctx, "tv3_gemv_noalloc", plan, trans, x, y_na)
    assert rc == 0, tval3_amd.lib().tv3_ctx_error(ctx.h).decode()
    ctx.sync()
    a, b = y_ref.cpu().numpy(), y_na.cpu().numpy()
    assert np.isfinite(a).all()
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    return x.cpu().numpy(), b


@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("trans", [0, 1])
def test_noalloc_dense_two_calls(ctx, bf16, trans):
    # first the larger operator (the A^T x partials fill 16 * 4096 doubles),
    # then a smaller one with ragged sizes
    for (m, n), seed in (((1229, 4096), 1), ((301, 1000), 2)):
        A = synth.gaussian_A(m, n, seed=seed)
        plan = ctx.plan_dense(A, bf16=bf16)
        x, y = _pair(ctx, plan, trans, seed)
        Ad = A.astype(np.float64)
        if bf16:
            import torch
            Ad = torch.from_numpy(A).to(torch.bfloat16).to(torch.float64).numpy()
        want = (Ad.T @ x.astype(np.float64)) if trans else (Ad @ x.astype(np.float64))
        absrow = (np.abs(Ad.T) @ np.abs(x.astype(np.float64))) if trans else (np.abs(Ad) @ np.abs(x.astype(np.float64)))
        assert np.all(np.abs(y - want) <= 2.4e-7 * absrow + 1e-30)
        plan.close()


@pytest.mark.parametrize("trans", [0, 1])
def test_noalloc_geometry_and_sparse_two_calls(ctx, trans):
    for K, grid, seed in ((128, 128, 3), (64, 48, 4)):
        geo = synth.ring_geometry(K, grid)
        gp = ctx.plan_geometry(geo)
        _pair(ctx, gp, trans, seed)
        ptr, lin = gp.walk()
        csr = synth.csr_from_walk(ptr, lin, geo.n)
        sp = ctx.plan_sparse(geo.m, geo.n, *csr)
        _pair(ctx, sp, trans, seed + 10)
        sp.close()
        gp.close()


def test_noalloc_rejects_dense_transpose_beyond_its_scratch(ctx):
    n = 65536 + 8
    A = np.ones((2, n), np.float32)
    plan = ctx.plan_dense(A)
    t = ctx.torch
    x = t.ones(2, dtype=t.float32, device=ctx.device)
    y = t.zeros(n, dtype=t.float32, device=ctx.device)
    rc = _call(ctx, "tv3_gemv_noalloc", plan, 1, x, y)
    assert rc == 1
    assert "65536" in tval3_amd.lib().tv3_ctx_error(ctx.h).decode()
    # A x of the same plan needs no partials
    x2 = t.ones(n, dtype=t.float32, device=ctx.device)
    y2 = t.zeros(2, dtype=t.float32, device=ctx.device)
    assert _call(ctx, "tv3_gemv_noalloc", plan, 0, x2, y2) == 0
    ctx.sync()
    assert y2.cpu().numpy().tolist() == [float(n), float(n)]
    plan.close()
