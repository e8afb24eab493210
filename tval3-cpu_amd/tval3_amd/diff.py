"""Differentiation harness: K unrolled TVAL3 inner iterations as a
torch.autograd.Function over the C-ABI (tv3_unrolled_forward / _backward).

The map (see include/tval3_amd.h): K inner iterations of the first outer
iteration of the reference solver (2d/tval3_cpu_2d.cpp:752-830) with no
stopping test and the line-search decisions of a forward solve held fixed.
Gradients flow to the measurements b and to the scalar mu; the adjoint is the
library's hand-written kernels, with checkpoints every `ckpt` steps.
"""
import numpy as np
import torch

from .abi import default_options


def decisions(trace, K):
    """(alpha[K], first[K]) from the trace of a solve with maxit = K."""
    if len(trace) < K:
        raise ValueError(f"trace has {len(trace)} records, need {K}")
    alpha = np.ones(K)
    first = np.zeros(K, np.int32)
    for k, t in enumerate(trace[:K]):
        if t.outer != 0:
            raise ValueError("the unrolled map covers the first outer iteration only")
        if k == 0 or t.fallback:
            first[k] = 1
            alpha[k] = 1.0
        else:
            alpha[k] = t.alpha
    return alpha, first


def options_for(mu, **kw):
    """Options of the unrolled map: no continuation, no stopping test."""
    kw.setdefault("scal_double", 1)
    kw.setdefault("ls_exact", 1)
    return default_options(mu=float(mu), mu0=float(mu), tol_inn=0.0, tol=0.0, **kw)


def record(ctx, plan, b, dims, mu, K, **kw):
    """Runs the solver for K iterations and returns its decisions."""
    o = options_for(mu, maxit=K, **kw)
    _, info, tr = ctx.solve(plan, b, (dims[0], dims[1], 1), o, trace_cap=K)
    return decisions(tr, K)


class _Unrolled(torch.autograd.Function):
    @staticmethod
    def forward(fctx, b, mu, tctx, plan, dims, K, ckpt, alpha, first, kw):
        o = options_for(float(mu), maxit=K, **kw)
        U, tape = tctx.unrolled_forward(plan, b.detach(), dims, o, K, ckpt, alpha, first)
        fctx.tape = tape
        fctx.tctx = tctx
        return U

    @staticmethod
    def backward(fctx, dU):
        db, dmu = fctx.tctx.unrolled_backward(fctx.tape, dU.contiguous())
        fctx.tape.close()
        return (db, torch.tensor(dmu, dtype=torch.float64, device=db.device),
                None, None, None, None, None, None, None, None)


def unrolled(tctx, plan, b, mu, dims, K=20, ckpt=5, alpha=None, first=None, **kw):
    """U_K of the unrolled map, differentiable in b (device float32 tensor)
    and mu (0-d float64 tensor).  Decisions are recorded from a solve at the
    given b and mu unless alpha / first are passed."""
    if alpha is None or first is None:
        alpha, first = record(tctx, plan, b.detach(), dims, float(mu), K, **kw)
    return _Unrolled.apply(b, mu, tctx, plan, dims, K, ckpt, alpha, first, kw)
