"""Degenerate problems on the GPU against the oracle.

The reference's scaleb (2d/tval3_cpu_2d.cpp:228-266) divides by
b_dif = |b[imax] - b[imin]| unguarded, so b = 0 or a b of constant magnitude
gives scl = inf and 0 * inf = NaN; an all-zero A gives U = 0, s = y = 0 and
tau = 0 / 0 (2d:449-450).  The engine keeps that arithmetic: U comes back all
NaN, rel_chg is NaN (the flag a caller tests), the inner loop runs to maxit
because every NaN comparison fails, and nothing hangs.  The iteration counts
and each iteration's backtracking and fallback decisions must be the
oracle's, which pins the sense of every comparison on the NaN path.

A zero column of A (a pixel no ray sees) is not degenerate: it is solved to
the ordinary bar.  In a 64-instance batch, lanes with NaN problems leave the
other lanes bit-identical to the same batch with ordinary problems there.
"""
import math

import numpy as np
import pytest

import tval3_amd
from tval3_amd import synth
import tv3_oracle as O

pytestmark = pytest.mark.gpu


def _traj(tr):
    return [(t.iter, t.outer, t.backtracks, t.fallback) for t in tr]


def _dense(q, m, seed):
    A = synth.gaussian_A(m, q * q, seed)
    x = synth.ellipse_phantom(q, q, 8, seed)
    return A, synth.add_noise(A @ x, 0.01, seed)


def _degenerate_cases():
    q, m = 24, 288
    A, b = _dense(q, m, 3)
    alt = np.where(np.arange(m) % 3 == 0, -0.75, 0.75).astype(np.float32)
    return [("b = 0", A, np.zeros(m, np.float32)),
            ("b constant", A, np.full(m, 2.5, np.float32)),
            ("|b| constant", A, alt),
            ("A = 0", np.zeros_like(A), b)], (q, q, 1)


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("ls_exact", [0, 1])
def test_nan_problems_follow_the_oracle(ctx, f64, ls_exact):
    cases, dims = _degenerate_cases()
    opts = tval3_amd.default_options(maxit=60, ls_exact=ls_exact, scal_double=int(f64))
    for name, A, b in cases:
        print(f"case {name} f64={f64} ls_exact={ls_exact}", flush=True)
        if f64:
            plan = ctx.plan_dense(A.astype(np.float64), f64=True)
            U, info, tr = ctx.solve(plan, b.astype(np.float64), dims, opts, trace_cap=64, f64=True)
            Uo, info_o, tro = O.solve_f64(O.Op.dense(A), b.astype(np.float64), dims, opts,
                                          trace_cap=64)
        else:
            plan = ctx.plan_dense(A)
            U, info, tr = ctx.solve(plan, b, dims, opts, trace_cap=64)
            Uo, info_o, tro = O.solve(O.Op.dense(A), b, dims, opts, trace_cap=64)
        Ug = U.cpu().numpy()
        assert np.isnan(Uo).all(), name
        assert np.isnan(Ug).all(), name
        assert math.isnan(info.rel_chg) and math.isnan(info_o.rel_chg), name
        assert info.total_iters == info_o.total_iters == opts.maxit, name
        assert info.outer_iters == info_o.outer_iters, name
        assert _traj(tr) == _traj(tro), name


def test_zero_columns_are_ordinary(ctx):
    q, m = 24, 288
    A, _ = _dense(q, m, 5)
    A[:, [0, 7, 300, q * q - 1]] = 0.0
    x = synth.ellipse_phantom(q, q, 8, 5)
    b = synth.add_noise(A @ x, 0.01, 5)
    dims = (q, q, 1)
    opts = tval3_amd.default_options(maxit=200, ls_exact=1)
    U, info, tr = ctx.solve(ctx.plan_dense(A), b, dims, opts, trace_cap=256)
    Uo, info_o, tro = O.solve(O.Op.dense(A), b, dims, opts, trace_cap=256)
    Ug = U.cpu().numpy()
    assert np.isfinite(Ug).all()
    rel = np.linalg.norm(Ug.astype(np.float64) - Uo) / np.linalg.norm(Uo)
    assert rel <= 1e-4, rel
    assert info.total_iters == info_o.total_iters
    assert _traj(tr) == _traj(tro)


def test_nan_lanes_leave_the_batch_alone(ctx):
    geo = synth.ring_geometry(32, 32)
    plan = ctx.plan_geometry(geo)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(32, seed=4)
    y = op.gemv(x.ravel())
    B = 64
    good = [synth.add_noise(y, 0.01, seed=100 + k) for k in range(B)]
    bad = list(good)
    nan_lanes = (0, 9, 31, 63)
    for k in nan_lanes:
        bad[k] = np.zeros_like(good[k]) if k % 2 else np.full_like(good[k], 0.5)
    opts = tval3_amd.default_options(maxit=120)
    dims = [(32, 32, 1)] * B
    Ug, ig = ctx.solve_batch([plan] * B, good, dims, opts)
    Ub, ib = ctx.solve_batch([plan] * B, bad, dims, opts)
    for k in range(B):
        ub = Ub[k].cpu().numpy()
        if k in nan_lanes:
            assert np.isnan(ub).all(), k
            assert math.isnan(ib[k].rel_chg) and ib[k].total_iters == opts.maxit, k
        else:
            assert np.array_equal(ub, Ug[k].cpu().numpy()), k
            assert ib[k].total_iters == ig[k].total_iters, k
            assert ib[k].outer_iters == ig[k].outer_iters, k
    Uo, info_o, _ = O.solve(op, bad[9], (32, 32, 1), opts)
    assert np.isnan(Uo).all() and info_o.total_iters == ib[9].total_iters
