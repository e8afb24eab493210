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
    """Relative L2 distance in FP64 (complex arrays in complex128)."""
    t = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = np.asarray(a, t)
    b = np.asarray(b, t)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def spread(op, b, dims, opts, Uo, solve=None):
    """Max relative change of the oracle's double / complex-double U under
    other reduction orders and perturbed b (fixed seed).  solve: the
    oracle's solve for b's type (default O.solve_f64; O.solve_c for complex
    b)."""
    solve = solve or O.solve_f64
    b = np.asarray(b)
    b = b.astype(np.complex128 if np.iscomplexobj(b) else np.float64)
    s = 0.0
    try:
        for chunk in (1021, 257, 64):
            O.set_sum_chunk(chunk)
            Up, _, _ = solve(op, b, dims, opts)
            s = max(s, rel(Up, Uo))
    finally:
        O.set_sum_chunk(4096)
    rng = np.random.default_rng(1234)
    for eps in (1e-16, 1e-15):
        bp = b * (1 + eps * rng.standard_normal(b.size))
        Up, _, _ = solve(op, bp, dims, opts)
        s = max(s, rel(Up, Uo))
    return s


def bound(op, b, dims, opts, Uo, solve=None):
    """The accepted distance from the oracle: 10x its own spread, at least
    1e-12."""
    return max(10 * spread(op, b, dims, opts, Uo, solve), 1e-12)
