"""Agreement bar for double/double solves (imported by the FP64 tests).

Two FP64 solves that sum in different orders agree to rounding only, and the
TVAL3 iteration amplifies rounding differences (about tenfold every five
inner iterations on the 64x64 ring).  The bar is therefore relative to the
oracle's own sensitivity to the same kind of change: its result under other
summation orders of its FP64 reductions (chunk lengths 1021, 257, 64 instead
of 4096) and under rounding-level perturbations of b (1e-16, 1e-15).
"""
import numpy as np

import tv3_oracle as O


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def spread(op, b, dims, opts, Uo):
    """Max relative change of the oracle's double/double U under other
    reduction orders and perturbed b (fixed seed)."""
    b = np.asarray(b, np.float64)
    s = 0.0
    try:
        for chunk in (1021, 257, 64):
            O.set_sum_chunk(chunk)
            Up, _, _ = O.solve_f64(op, b, dims, opts)
            s = max(s, rel(Up, Uo))
    finally:
        O.set_sum_chunk(4096)
    rng = np.random.default_rng(1234)
    for eps in (1e-16, 1e-15):
        bp = b * (1 + eps * rng.standard_normal(b.size))
        Up, _, _ = O.solve_f64(op, bp, dims, opts)
        s = max(s, rel(Up, Uo))
    return s


def bound(op, b, dims, opts, Uo):
    """The accepted distance from the oracle: 10x its own spread, at least
    1e-12."""
    return max(10 * spread(op, b, dims, opts, Uo), 1e-12)
