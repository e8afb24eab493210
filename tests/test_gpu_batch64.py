"""64-lane groups of instances sharing one projector, against the oracle.

Every lane of a full-size ring batch at maxit = 0 (U is the scaled A^T b: one
unmasked transpose per group, whose lists end in partial rounds of four), and
sampled lanes of whole solves, 2-D and 3-D, whose masked launches mix full,
lane-compacted and W = 1 products.  The bar is bit-exact: the GPU sums the
same float products in FP64 and rounds once, as the oracle does.
"""
import numpy as np
import pytest

import tval3_amd
from tval3_amd import synth
import tv3_oracle as O

pytestmark = pytest.mark.gpu


def _ring_batch(K, grid, B, seed0):
    geo = synth.ring_geometry(K, grid)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(grid, seed=seed0)
    y = op.gemv(x)
    bs = [synth.add_noise(y, 0.005, seed0 + i) for i in range(B)]
    return geo, op, x, bs


def test_full_size_ring_atb_every_lane(ctx):
    """256 transducers on 256x256, 130 instances (three groups, the last
    padded), maxit = 0: every lane's U equals the oracle's."""
    geo, op, x, bs = _ring_batch(256, 256, 130, 7)
    plan = ctx.plan_geometry(geo)
    opts = tval3_amd.default_options(maxit=0)
    Us, _ = ctx.solve_batch([plan] * len(bs), bs, [(256, 256, 1)] * len(bs), opts)
    for i, b in enumerate(bs):
        Uo, _, _ = O.solve(op, b, (256, 256, 1), opts)
        assert np.array_equal(Us[i].cpu().numpy(), Uo), i


def test_ring_batch_solves(ctx):
    """130 instances of the 48x48 ring, 60 iterations with the reference's
    backtracking: sampled lanes equal the oracle bit for bit, with the same
    iteration counts."""
    geo, op, x, bs = _ring_batch(48, 48, 130, 100)
    plan = ctx.plan_geometry(geo)
    opts = tval3_amd.default_options(maxit=60, ls_exact=1)
    Us, infos = ctx.solve_batch([plan] * len(bs), bs, [(48, 48, 1)] * len(bs), opts, Uts=[x] * len(bs))
    for i in (0, 1, 17, 63, 64, 100, 129):
        Uo, info_o, _ = O.solve(op, bs[i], (48, 48, 1), opts, Ut=x)
        assert np.array_equal(Us[i].cpu().numpy(), Uo), i
        assert infos[i].total_iters == info_o.total_iters, i


def test_hemisphere_batch_solves(ctx):
    """A shared 3-D geometry (64 transducers on a 16^3 hemisphere) in two
    64-lane groups: sampled lanes equal the oracle bit for bit."""
    grid = 16
    geo = synth.hemisphere_geometry(64, grid)
    op = O.Op.geometry(geo)
    x = synth.ellipsoid_phantom(grid, 6, seed=5)
    y = op.gemv(x)
    B = 70
    bs = [synth.add_noise(y, 0.005, 20 + i) for i in range(B)]
    plan = ctx.plan_geometry(geo)
    dims = (grid, grid, grid)
    opts = tval3_amd.default_options(maxit=30, ls_exact=1)
    Us, infos = ctx.solve_batch([plan] * B, bs, [dims] * B, opts, Uts=[x] * B)
    for i in (0, 5, 63, 64, 69):
        Uo, info_o, _ = O.solve(op, bs[i], dims, opts, Ut=x)
        assert np.array_equal(Us[i].cpu().numpy(), Uo), i
        assert infos[i].total_iters == info_o.total_iters, i


@pytest.mark.parametrize("kw", [dict(ls_exact=0), dict(ls_exact=1),
                                dict(ls_exact=0, nonneg=1, scal_double=1),
                                dict(ls_exact=1, tv_iso=1)])
def test_four_instances_per_lane_passes(ctx, kw, monkeypatch):
    """The 2-D float vector passes of 64-lane groups with four instances per
    lane (16-byte accesses) against one instance per lane: every lane of a
    130-instance ring batch (three groups, the last padded, lanes finishing
    at different iterations) gives the same bits and iteration counts, and
    sampled lanes agree with the oracle."""
    geo, op, x, bs = _ring_batch(48, 48, 130, 300)
    plan = ctx.plan_geometry(geo)
    opts = tval3_amd.default_options(maxit=80, **kw)
    res = {}
    for v4 in ("0", "1"):
        monkeypatch.setenv("TV3_VEC4", v4)
        Us, infos = ctx.solve_batch([plan] * len(bs), bs, [(48, 48, 1)] * len(bs), opts)
        res[v4] = (np.stack([u.cpu().numpy() for u in Us]), [i.total_iters for i in infos])
    assert res["0"][1] == res["1"][1]
    for k in range(len(bs)):
        assert np.array_equal(res["0"][0][k], res["1"][0][k]), k
    for i in (0, 77, 129):
        Uo, info_o, _ = O.solve(op, bs[i], (48, 48, 1), opts)
        if kw["ls_exact"]:
            assert np.array_equal(res["1"][0][i], Uo), i
            assert res["1"][1][i] == info_o.total_iters, i
        else:
            # the quadratic line search (ls_exact = 0) is held to the
            # tolerance the single-instance tests give it
            rel = np.linalg.norm(res["1"][0][i].astype(np.float64) - Uo) / np.linalg.norm(Uo)
            assert rel <= 1e-4, (i, rel)


def test_finished_instances_in_four_instance_passes(ctx, monkeypatch):
    """Instances that finish tens of iterations apart (noise from 0 to 5%):
    once an instance has finished, the four-instances-per-lane passes
    overwrite its lane of the working arrays.  Every lane's result and
    iteration count equals the one-instance-per-lane kernels', and sampled
    early and late finishers equal the oracle."""
    geo = synth.ring_geometry(48, 48)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(48, seed=301)
    y = op.gemv(x)
    B = 130
    bs = [synth.add_noise(y, 0.05 * (i % 13) / 12, 400 + i) if i % 13 else y.copy() for i in range(B)]
    plan = ctx.plan_geometry(geo)
    opts = tval3_amd.default_options(maxit=150, tol=1e-3, tol_inn=1e-3, ls_exact=1)
    res = {}
    for v4 in ("0", "1"):
        monkeypatch.setenv("TV3_VEC4", v4)
        Us, infos = ctx.solve_batch([plan] * B, bs, [(48, 48, 1)] * B, opts)
        res[v4] = (np.stack([u.cpu().numpy() for u in Us]), [i.total_iters for i in infos])
    iters = res["1"][1]
    assert res["0"][1] == iters
    assert max(iters) - min(iters) >= 30, (min(iters), max(iters))
    for k in range(B):
        assert np.array_equal(res["0"][0][k], res["1"][0][k]), k
    early, late = int(np.argmin(iters)), int(np.argmax(iters))
    for i in sorted({early, late, 64, 129}):
        Uo, info_o, _ = O.solve(op, bs[i], (48, 48, 1), opts)
        assert np.array_equal(res["1"][0][i], Uo), i
        assert iters[i] == info_o.total_iters, i
