"""Differentiation through K unrolled TVAL3 iterations (C-ABI
tv3_unrolled_forward / tv3_unrolled_backward, tval3_amd.diff).

CPU: the FP64 reference map (tests/diff_ref.py) reproduces the oracle solver's
K-th iterate, and its autograd agrees with central finite differences.
GPU: the unrolled forward equals the solver's K-th iterate; the hand-written
adjoint agrees with the FP64 autograd reference and with central finite
differences on 32 seeded directions (h = 1e-3); checkpoint spacing does not
change the result.  Loss = |U_K - x_true|^2, K = 20, checkpoints every 5, ring
of 128 transducers on 128x128 (8,128 rays).
"""
import numpy as np
import pytest
import torch

import tv3_oracle as O
from tval3_amd import diff
import diff_ref as R

K = 20


def _oracle_decisions(op, b, dims, mu, K):
    o = diff.options_for(mu, maxit=K)
    U, info, tr = O.solve(op, b, dims, o, trace_cap=K)
    return U, diff.decisions(tr, K)


def _ref_loss_and_grads(A, AT, b, mu, beta, q, p, alpha, first, x):
    bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    mt = torch.tensor(mu, dtype=torch.float64, requires_grad=True)
    U = R.unrolled(A, AT, bt, mt, beta, q, p, len(alpha), alpha, first)
    L = ((U - torch.from_numpy(x.astype(np.float64))) ** 2).sum()
    L.backward()
    return float(L.detach()), bt.grad.numpy(), float(mt.grad), U.detach().numpy()


def _ref_loss(A, AT, b, mu, beta, q, p, alpha, first, x):
    with torch.no_grad():
        U = R.unrolled(A, AT, torch.from_numpy(b), torch.tensor(mu, dtype=torch.float64),
                       beta, q, p, len(alpha), alpha, first)
        return float(((U - torch.from_numpy(x.astype(np.float64))) ** 2).sum())


@pytest.fixture(scope="module")
def small():
    geo, op, x, b, A, AT = R.c4_problem(48, 48, seed=3)
    return geo, op, x, b.astype(np.float64), A, AT


def test_reference_map_is_the_solver(small):
    geo, op, x, b, A, AT = small
    q = p = 48
    mu, beta = 256.0, 32.0
    Uo, (alpha, first) = _oracle_decisions(op, b.astype(np.float32), (q, p, 1), mu, K)
    _, _, _, Ur = _ref_loss_and_grads(A, AT, b.astype(np.float32).astype(np.float64), mu, beta,
                                      q, p, alpha, first, x)
    assert np.linalg.norm(Ur - Uo) / np.linalg.norm(Uo) < 1e-4


def test_reference_autograd_matches_finite_differences(small):
    geo, op, x, b, A, AT = small
    q = p = 48
    mu, beta = 256.0, 32.0
    _, (alpha, first) = _oracle_decisions(op, b.astype(np.float32), (q, p, 1), mu, K)
    L, gb, gmu, _ = _ref_loss_and_grads(A, AT, b, mu, beta, q, p, alpha, first, x)
    rng = np.random.default_rng(0)
    h = 1e-3
    for _ in range(6):
        v = rng.standard_normal(b.size)
        v /= np.linalg.norm(v)
        fd = (_ref_loss(A, AT, b + h * v, mu, beta, q, p, alpha, first, x)
              - _ref_loss(A, AT, b - h * v, mu, beta, q, p, alpha, first, x)) / (2 * h)
        # the map is piecewise smooth (shrinkage kinks): an absolute floor
        # proportional to |grad| covers directions whose derivative is small
        assert fd == pytest.approx(gb @ v, rel=1e-4, abs=1e-3 * np.linalg.norm(gb))
    hm = 1e-4 * mu  # mu steps of 1e-3 mu cross shrinkage kinks on this problem
    fdm = (_ref_loss(A, AT, b, mu + hm, beta, q, p, alpha, first, x)
           - _ref_loss(A, AT, b, mu - hm, beta, q, p, alpha, first, x)) / (2 * hm)
    assert fdm == pytest.approx(gmu, rel=1e-4, abs=1e-9)


# ---- GPU ----------------------------------------------------------------------
@pytest.fixture(scope="module")
def c4(ctx):
    geo, op, x, b, A, AT = R.c4_problem(128, 128, seed=0)
    plan = ctx.plan_geometry(geo)
    return geo, plan, x, b, A, AT


def _gpu_loss_grads(ctx, plan, b, mu, x, alpha, first, ckpt=5):
    bt = ctx.tensor(b).requires_grad_(True)
    mt = torch.tensor(mu, dtype=torch.float64, device=bt.device, requires_grad=True)
    U = diff.unrolled(ctx, plan, bt, mt, (128, 128), K=len(alpha), ckpt=min(ckpt, len(alpha)),
                      alpha=alpha, first=first)
    L = ((U.double() - ctx.tensor(x).double()) ** 2).sum()
    L.backward()
    return float(L.detach()), bt.grad.double().cpu().numpy(), float(mt.grad), U.detach().cpu().numpy()


@pytest.mark.gpu
def test_unrolled_forward_is_the_solver(ctx, c4):
    geo, plan, x, b, A, AT = c4
    mu = 256.0
    print("diff case: unrolled forward vs solver, ring 128, K=20", flush=True)
    alpha, first = diff.record(ctx, plan, ctx.tensor(b), (128, 128), mu, K)
    o = diff.options_for(mu, maxit=K)
    Us, info, _ = ctx.solve(plan, b, (128, 128, 1), o)
    U, tape = ctx.unrolled_forward(plan, b, (128, 128), o, K, 5, alpha, first)
    tape.close()
    Us, U = Us.cpu().numpy(), U.cpu().numpy()
    assert info.total_iters == K
    # the solver rounds its dot products to float (as cblas_sdot does); the
    # unrolled map keeps FP64 scalars.  tau differs by ~1e-7 per step and the
    # difference grows along the trajectory (8e-5 at K = 20 on this problem)
    assert np.linalg.norm(U - Us) / np.linalg.norm(Us) < 5e-4


# FP32 forward drift from the FP64 map (measured with the FP32 oracle on this
# problem: U 1.1e-5 at K = 5, 7.1e-4 at K = 20) bounds the agreement of the
# FP32 adjoint with FP64 autograd; the finite-difference check measures the
# error of each directional derivative against |grad|, where a wrong adjoint
# term shows at the 1e-2 level and FP32 noise at 1e-4.  dL/dmu is a sum of
# per-use terms that cancel: a first step's tau * d is nearly independent of
# mu, so its two terms (about 2e5 each here) cancel to O(1); its error is
# bounded against the sum of |terms|, where FP32 resolves about 1e-8 and a
# missing or wrong term shows at 1e-2 and above.
@pytest.mark.gpu
@pytest.mark.parametrize("Kc,tol", [(5, 1e-3), (20, 2e-2)])
def test_adjoint_matches_fp64_autograd_and_fd(ctx, c4, Kc, tol):
    geo, plan, x, b, A, AT = c4
    mu, beta = 256.0, 32.0
    print(f"diff case: adjoint vs FP64 autograd + 32 FD directions, ring 128, K={Kc}", flush=True)
    alpha, first = diff.record(ctx, plan, ctx.tensor(b), (128, 128), mu, Kc)
    Lg, gb, gmu, Ug = _gpu_loss_grads(ctx, plan, b, mu, x, alpha, first)
    b64 = b.astype(np.float64)
    Lr, gbr, gmur, Ur = _ref_loss_and_grads(A, AT, b64, mu, beta, 128, 128, alpha, first, x)
    rel_b = np.linalg.norm(gb - gbr) / np.linalg.norm(gbr)
    print(f"K={Kc}: loss gpu {Lg:.6e} ref {Lr:.6e}; dL/db rel L2 {rel_b:.2e}; "
          f"dL/dmu gpu {gmu:.6e} ref {gmur:.6e}")
    assert Lg == pytest.approx(Lr, rel=tol / 4)
    assert rel_b < tol
    scale = R.mu_term_scale(A, AT, b64, mu, beta, 128, 128, alpha, first, x)
    print(f"K={Kc}: |dmu gpu - ref| / sum|terms| = {abs(gmu - gmur) / scale:.2e} "
          f"(sum|terms| {scale:.3e})")
    assert abs(gmu - gmur) <= 1e-6 * scale
    # central finite differences (h = 1e-3) of the FP64 map on 32 seeded unit
    # directions against the GPU adjoint's directional derivatives
    worst = 0.0
    for s in range(32):
        v = np.random.default_rng(s).standard_normal(b.size)
        v /= np.linalg.norm(v)
        fd = (_ref_loss(A, AT, b64 + 1e-3 * v, mu, beta, 128, 128, alpha, first, x)
              - _ref_loss(A, AT, b64 - 1e-3 * v, mu, beta, 128, 128, alpha, first, x)) / 2e-3
        worst = max(worst, abs(fd - gb @ v) / np.linalg.norm(gbr))
    print(f"K={Kc}: worst |FD - <grad, v>| / |grad| over 32 directions {worst:.2e}")
    assert worst < 1e-3


@pytest.mark.gpu
def test_checkpoint_spacing_does_not_change_gradients(ctx, c4):
    geo, plan, x, b, A, AT = c4
    mu = 256.0
    print("diff case: checkpoint spacing 1 / 5 / 20", flush=True)
    alpha, first = diff.record(ctx, plan, ctx.tensor(b), (128, 128), mu, K)
    res = [_gpu_loss_grads(ctx, plan, b, mu, x, alpha, first, ckpt=c) for c in (1, 5, 20)]
    for r in res[1:]:
        np.testing.assert_array_equal(r[1], res[0][1])
        assert r[2] == res[0][2]
