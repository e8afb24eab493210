"""Pins the CPU oracle (runs without a GPU).

The reference ships no tests or golden vectors and cannot be built here (it
needs four project headers it does not ship and Intel MKL headers), so each
oracle function is pinned by closed forms, exact linear-algebra identities,
brute force on tiny inputs and an independent transcription of the Bresenham
recurrence (tests/golden).
"""
import json
import os

import numpy as np
import pytest

from tval3_amd import synth
from tval3_amd.abi import default_options
import tv3_oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------
# Bresenham walk (3d/mat_mult.h:241-359)
# ---------------------------------------------------------------------------
def test_walk_matches_golden():
    g = json.load(open(os.path.join(GOLDEN, "bresenham_walks.json")))
    grid = tuple(g["grid"])
    segs = g["segments"]
    K = len(segs)
    xe = [s["start"][0] for s in segs]; ye = [s["start"][1] for s in segs]
    ze = [s["start"][2] for s in segs]
    xr = [s["end"][0] for s in segs]; yr = [s["end"][1] for s in segs]
    zr = [s["end"][2] for s in segs]
    # emitter k pairs only with receiver k
    up = np.zeros((K, K), np.int32)
    up[np.arange(K), np.arange(K)] = np.arange(1, K + 1)
    geo = synth.GeometryArrays(1, up.ravel(), xe, ye, ze, xr, yr, zr, grid, K)
    ptr, lin = O.Op.geometry(geo).walk()
    for k, s in enumerate(segs):
        assert list(lin[ptr[k]:ptr[k + 1]]) == s["lin"], k


def _coords(lin, grid):
    q, p, _ = grid
    return lin % q, (lin // q) % p, lin // (q * p)


@pytest.mark.parametrize("geo_fn,grid", [
    (lambda: synth.ring_geometry(128, 128), (128, 128, 1)),
    (lambda: synth.hemisphere_geometry(64, 32), (32, 32, 32)),
])
def test_walk_invariants(geo_fn, grid):
    geo = geo_fn()
    ptr, lin = O.Op.geometry(geo).walk()
    K = geo.xe.size
    iu = np.triu_indices(K, 1)
    for row in range(0, geo.m, max(1, geo.m // 400)):
        e, r = iu[0][row], iu[1][row]
        seg = lin[ptr[row]:ptr[row + 1]]
        x, y, z = _coords(seg, grid)
        s = np.array([geo.xe[e], geo.ye[e], geo.ze[e]])
        t = np.array([geo.xr[r], geo.yr[r], geo.zr[r]])
        d = t - s
        assert (x[0], y[0], z[0]) == tuple(s) and (x[-1], y[-1], z[-1]) == tuple(t)
        assert seg.size == np.abs(d).max() + 1
        assert np.unique(seg).size == seg.size
        steps = np.stack([np.diff(x), np.diff(y), np.diff(z)])
        assert np.all(np.abs(steps) <= 1)
        major = int(np.argmax(np.abs(d)))   # first maximum: ties go to x, then y
        if d[major] != 0:
            assert np.all(steps[major] == np.sign(d[major]))


def test_ring_nnz_counts():
    assert O.Op.geometry(synth.ring_geometry(256, 256)).nnz == 4821527
    assert O.Op.geometry(synth.ring_geometry(128, 128)).nnz == 604831


def _explicit(geo):
    ptr, lin = O.Op.geometry(geo).walk()
    A = np.zeros((geo.m, geo.n))
    rows = np.repeat(np.arange(geo.m), np.diff(ptr))
    np.add.at(A, (rows, lin), np.float64(np.float32(geo.scale)))
    return A


def test_geometry_operator_brute_force():
    geo = synth.ring_geometry(24, 20, scale=0.5)
    op = O.Op.geometry(geo)
    A = _explicit(geo)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(geo.n).astype(np.float32)
    y = rng.standard_normal(geo.m).astype(np.float32)
    np.testing.assert_allclose(op.gemv(x), A @ x, rtol=3e-7, atol=1e-6)
    np.testing.assert_allclose(op.gemv(y, trans=True), A.T @ y, rtol=3e-7, atol=1e-6)


def test_dense_and_sparse_operators():
    A = synth.gaussian_A(50, 70, 1)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(70).astype(np.float32)
    y = rng.standard_normal(50).astype(np.float32)
    Ad = A.astype(np.float64)
    op = O.Op.dense(A)
    # FP64 accumulation rounded once: within half an ulp of the exact value
    np.testing.assert_allclose(op.gemv(x), Ad @ x, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(op.gemv(y, True), Ad.T @ y, rtol=1e-6, atol=1e-6)
    A[np.abs(A) < 0.1] = 0
    import scipy.sparse as sp
    S = sp.csr_matrix(A)
    T = sp.csc_matrix(A)
    sop = O.Op.sparse(50, 70, S.indptr, S.indices, S.data, T.indptr, T.indices, T.data)
    np.testing.assert_allclose(sop.gemv(x), A.astype(np.float64) @ x, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(sop.gemv(y, True), A.T.astype(np.float64) @ y, rtol=1e-6, atol=1e-6)


# ---------------------------------------------------------------------------
# D / Dt (2d/tval3_cpu_2d.cpp:268-293, 3d/tval3_cpu_3d.cpp:365-425)
# ---------------------------------------------------------------------------
def _D_matrix(dims):
    n = int(np.prod(dims))
    z3 = dims[2] > 1
    M = np.zeros((3 * n if z3 else 2 * n, n))
    for j in range(n):
        e = np.zeros(n, np.float32); e[j] = 1
        Ux, Uy, Uz = O.D(e, dims)
        M[:n, j] = Ux; M[n:2 * n, j] = Uy
        if z3:
            M[2 * n:, j] = Uz
    return M


def _Dt_matrix(dims, exact):
    n = int(np.prod(dims))
    z3 = dims[2] > 1
    k = 3 if z3 else 2
    M = np.zeros((n, k * n))
    z = np.zeros(n, np.float32)
    for j in range(k * n):
        X, Y, Z = z.copy(), z.copy(), z.copy()
        [X, Y, Z][j // n][j % n] = 1
        M[:, j] = O.Dt(X, Y, Z, dims, exact)
    return M


@pytest.mark.parametrize("dims", [(5, 4, 1), (6, 3, 1), (4, 3, 3), (3, 5, 2)])
def test_D_closed_form(dims):
    q, p, r = dims
    U = np.random.default_rng(0).standard_normal(q * p * r).astype(np.float32)
    V = U.reshape(r, p, q)
    Ux, Uy, Uz = O.D(U, dims)
    np.testing.assert_array_equal(Ux, (np.roll(V, -1, 2) - V).ravel())
    np.testing.assert_array_equal(Uy, (np.roll(V, -1, 1) - V).ravel())
    if r > 1:
        np.testing.assert_array_equal(Uz, (np.roll(V, -1, 0) - V).ravel())


@pytest.mark.parametrize("dims", [(5, 4, 1), (7, 2, 1), (4, 3, 3), (3, 5, 2)])
def test_Dt_is_adjoint_except_reference_3d_defect(dims):
    q, p, r = dims
    n, s = q * p * r, q * p
    Dm = _D_matrix(dims)
    exact = _Dt_matrix(dims, True)
    np.testing.assert_array_equal(exact, Dm.T)
    ref = _Dt_matrix(dims, False)
    if r == 1:
        np.testing.assert_array_equal(ref, Dm.T)
        return
    # the reference's 3-D Dt: row z*s + k (z >= 1, k < dim_x) gains
    # +Y[(z-1)*s + (dim_y-1)*dim_x + k] - Y[z*s + k]
    defect = np.zeros_like(ref)
    for z in range(1, r):
        for k in range(q):
            i = z * s + k
            defect[i, n + (z - 1) * s + (p - 1) * q + k] += 1
            defect[i, n + i] -= 1
    np.testing.assert_array_equal(ref - Dm.T, defect)


def test_Dt_operation_order_2d():
    """Each entry is ((X[xp] - X[i]) + Y[yp]) - Y[i] with float rounding at
    every step, the order of the reference's passes."""
    q, p = 9, 7
    rng = np.random.default_rng(3)
    X = (rng.standard_normal(q * p) * 10 ** rng.uniform(-3, 3, q * p)).astype(np.float32)
    Y = (rng.standard_normal(q * p) * 10 ** rng.uniform(-3, 3, q * p)).astype(np.float32)
    Xm, Ym = X.reshape(p, q), Y.reshape(p, q)
    want = ((np.roll(Xm, 1, 1) - Xm) + np.roll(Ym, 1, 0)) - Ym
    np.testing.assert_array_equal(O.Dt(X, Y, None, (q, p, 1)), want.ravel())


# ---------------------------------------------------------------------------
# shrinkage (2d/tval3_cpu_2d.cpp:588-611)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("sd", [False, True])
def test_shrink_anisotropic_closed_form(sd):
    rng = np.random.default_rng(4)
    n = 4000
    Ux, Uy, sx, sy = (rng.standard_normal(n).astype(np.float32) * 0.1 for _ in range(4))
    beta = 32.0
    Wx, Wy, _, lam1 = O.shrink(Ux, Uy, None, sx, sy, None, (n, 1, 1), beta, False, sd)
    T = np.float64 if sd else np.float32
    for U, s, W in ((Ux, sx, Wx), (Uy, sy, Wy)):
        ub = (U.astype(T) - s.astype(T) / T(beta)).astype(np.float32)
        w = np.maximum(np.abs(ub).astype(T) - T(1) / T(beta), 0)
        np.testing.assert_array_equal(W, np.where(ub >= 0, w, -w).astype(np.float32))
    ub = [(U.astype(np.float64) - s / beta) for U, s in ((Ux, sx), (Uy, sy))]
    exact = sum(np.maximum(np.abs(u) - 1 / beta, 0).sum() for u in ub)
    assert lam1 == pytest.approx(exact, rel=1e-5)


def test_shrink_special_cases():
    rng = np.random.default_rng(5)
    n = 500
    Ux, Uy, Uz = (rng.standard_normal(n).astype(np.float32) for _ in range(3))
    z = np.zeros(n, np.float32)
    # sigma = 0, beta -> infinity: W -> U
    Wx, Wy, Wz, _ = O.shrink(Ux, Uy, Uz, z, z, z, (10, 10, 5), 1e12, False, True)
    np.testing.assert_allclose(Wx, Ux, rtol=1e-6, atol=1e-11)
    np.testing.assert_allclose(Wz, Uz, rtol=1e-6, atol=1e-11)
    # isotropic with one non-zero component equals the anisotropic result
    Wi, Wi2, _, li = O.shrink(Ux, z, None, z, z, None, (n, 1, 1), 2.0, True, True)
    Wa, _, _, la = O.shrink(Ux, z, None, z, z, None, (n, 1, 1), 2.0, False, True)
    np.testing.assert_allclose(Wi, Wa, rtol=1e-6, atol=1e-7)
    assert np.all(Wi2 == 0)
    assert li == pytest.approx(la, rel=1e-12)
    # isotropic closed form: w = max(|u| - 1/beta, 0), W = w u / |u|
    Wx, Wy, _, l1 = O.shrink(Ux, Uy, None, z, z, None, (n, 1, 1), 4.0, True, True)
    r = np.hypot(Ux.astype(np.float64), Uy.astype(np.float64))
    w = np.maximum(r - 0.25, 0)
    np.testing.assert_allclose(Wx, w * Ux / r, rtol=1e-6, atol=1e-7)
    assert l1 == pytest.approx(w.sum(), rel=1e-12)


# ---------------------------------------------------------------------------
# solver (2d/tval3_cpu_2d.cpp:752-886)
# ---------------------------------------------------------------------------
def _dense(q, p, frac, seed, noise=0.01):
    n = q * p
    m = int(round(frac * n))
    A = synth.gaussian_A(m, n, seed)
    x = synth.ellipse_phantom(p, q, 6, seed)
    return A, x, synth.add_noise(A @ x, noise, seed)


def _scaleb(b):
    """Closed form of scaleb (2d:228-266): scale to |b_max - b_min| in
    [0.5, 1.5] where b_min / b_max are the first entries of smallest / largest
    magnitude."""
    a = np.abs(b)
    d = abs(float(b[np.argmax(a)]) - float(b[np.argmin(a)]))
    return 0.5 / d if d < 0.5 else (1.5 / d if d > 1.5 else 1.0)


@pytest.mark.parametrize("amp", [1e-3, 1.0, 50.0])
def test_maxit_zero_is_scaled_Atb(amp):
    A, x, b = _dense(12, 10, 0.5, 2)
    b = (b * amp).astype(np.float32)
    opts = default_options(maxit=0, scal_double=1)
    U, info, _ = O.solve(O.Op.dense(A), b, (12, 10, 1), opts)
    s = _scaleb(b)
    want = (A.astype(np.float64).T @ (np.float32(s) * b).astype(np.float64)) / s
    np.testing.assert_allclose(U, want, rtol=2e-6, atol=1e-7 * np.abs(want).max())
    assert info.total_iters == 0 and info.outer_iters == 0


@pytest.mark.parametrize("kind", ["zero", "constant", "constant magnitude"])
def test_scaleb_degenerate_is_nan(kind):
    """b_dif = 0 in scaleb (2d:256-258): scl = 0.5 / 0 = inf, and the scaled
    b is 0 * inf = NaN (b = 0) or +-inf, whose A^T product mixes signs into
    NaN; the unscaled U = NaN / inf stays NaN.  With iterations the NaN
    comparisons all fail: no backtracking, no inner or outer stop, so the
    solve runs to maxit in one outer iteration."""
    A, x, b = _dense(12, 10, 0.5, 2)
    b = {"zero": np.zeros_like(b), "constant": np.full_like(b, 3.0),
         "constant magnitude": np.where(np.arange(b.size) % 2, 0.4, -0.4).astype(np.float32)}[kind]
    a = np.abs(b)
    assert b[np.argmax(a)] == b[np.argmin(a)]  # b_dif = 0
    for maxit in (0, 25):
        opts = default_options(maxit=maxit, scal_double=1)
        U, info, tr = O.solve(O.Op.dense(A), b, (12, 10, 1), opts, trace_cap=32)
        assert np.isnan(U).all()
        assert info.total_iters == maxit and info.outer_iters == (1 if maxit else 0)
        assert [(t.backtracks, t.fallback) for t in tr] == [(0, 0)] * maxit


@pytest.mark.parametrize("sd", [0, 1])
def test_objective_bookkeeping(sd):
    """f is maintained incrementally (update_W subtracts and re-adds its terms,
    update_mlp likewise); at the end of every inner iteration it must equal the
    full Lagrangian lam1 + beta/2 lam2 + mu/2 lam3 - lam4 - lam5."""
    A, x, b = _dense(20, 16, 0.4, 6)
    opts = default_options(maxit=120, mu0=8.0, beta0=4.0, scal_double=sd)
    U, info, tr = O.solve(O.Op.dense(A), b, (20, 16, 1), opts, trace_cap=120)
    assert len(tr) == info.total_iters > 10
    for t in tr:
        full = t.lam1 + t.beta / 2 * t.lam2 + t.mu / 2 * t.lam3 - t.lam4 - t.lam5
        assert t.f == pytest.approx(full, rel=2e-4, abs=1e-4 * abs(full) + 1e-6)
    # continuation multiplies mu and beta by rate_cnt per outer iteration
    mus = sorted({t.mu for t in tr})
    assert mus[0] == 8.0 and all(b_ / a_ == 2.0 for a_, b_ in zip(mus, mus[1:]))


def test_nonmonotone_reference_recurrence():
    """C_k = (gam Q_{k-1} C_{k-1} + f_k) / Q_k with Q_k = gam Q_{k-1} + 1 inside
    an outer iteration (2d:802), checked on the trace in FP64."""
    A, x, b = _dense(16, 16, 0.4, 8)
    opts = default_options(maxit=60, scal_double=1)
    _, info, tr = O.solve(O.Op.dense(A), b, (16, 16, 1), opts, trace_cap=60)
    gam = opts.gam
    Q, C = 1.0, None
    prev_outer = tr[0].outer
    for k, t in enumerate(tr):
        if t.outer != prev_outer or k == 0:
            Q, C = 1.0, None
            prev_outer = t.outer
        if t.fallback:
            C = None
            continue
        if C is not None:
            Qn = gam * Q + 1
            assert t.C == pytest.approx((gam * Q * C + t.f) / Qn, rel=1e-6)
            Q = Qn
        else:
            Q = gam * Q + 1
        C = t.C


def test_float_and_double_scalars_agree():
    A, x, b = _dense(24, 24, 0.35, 9)
    op = O.Op.dense(A)
    Uf, inf_f, _ = O.solve(op, b, (24, 24, 1), default_options(maxit=200, scal_double=0), Ut=x)
    Ud, inf_d, _ = O.solve(op, b, (24, 24, 1), default_options(maxit=200, scal_double=1), Ut=x)
    assert np.linalg.norm(Uf - Ud) / np.linalg.norm(Ud) < 2e-2
    assert inf_f.rel_error < 0.1 and inf_d.rel_error < 0.1


def test_noise_free_reconstruction_converges():
    """Piecewise-constant phantom, 60 % Gaussian measurements, no noise: TV
    recovery must be accurate."""
    A, x, b = _dense(24, 24, 0.6, 10, noise=0.0)
    opts = default_options(maxit=400, tol=1e-6, scal_double=1, tv_iso=1)
    U, info, _ = O.solve(O.Op.dense(A), b, (24, 24, 1), opts, Ut=x)
    assert info.rel_error < 0.02


def test_thread_count_does_not_change_results():
    A, x, b = _dense(20, 20, 0.4, 12)
    op = O.Op.dense(A)
    opts = default_options(maxit=60)
    t0 = O.get_threads()
    try:
        O.set_threads(1)
        U1, _, _ = O.solve(op, b, (20, 20, 1), opts)
        O.set_threads(max(2, t0))
        U2, _, _ = O.solve(op, b, (20, 20, 1), opts)
    finally:
        O.set_threads(t0)
    np.testing.assert_array_equal(U1, U2)


def test_fp64_solve_close_to_fp32():
    """The double/double solve (FP64 reference) and the float/double solve of
    the same problem follow the same trajectory until a backtracking decision
    first differs (float rounding at an Armijo near-tie); over the first 20
    iterations they agree to float accuracy.  The FP64 gemv is the exact
    product."""
    A, x, b = _dense(20, 20, 0.5, 13)
    op = O.Op.dense(A)
    xd = np.random.default_rng(0).standard_normal(400)
    np.testing.assert_allclose(op.gemv_f64(xd), A.astype(np.float64) @ xd, rtol=1e-12, atol=1e-13)
    opts = default_options(maxit=20, scal_double=1)
    U64, i64, t64 = O.solve_f64(op, b, (20, 20, 1), opts, Ut=x, trace_cap=20)
    U32, i32, t32 = O.solve(op, b, (20, 20, 1), opts, Ut=x, trace_cap=20)
    assert [t.backtracks for t in t64] == [t.backtracks for t in t32]
    assert np.linalg.norm(U32 - U64) / np.linalg.norm(U64) < 1e-4
    opts = default_options(maxit=300, scal_double=1)
    U64, i64, _ = O.solve_f64(op, b, (20, 20, 1), opts, Ut=x)
    assert i64.rel_error < 0.05


def test_f64_bar_separates_reduction_order_from_float_rounding():
    """The FP64 agreement bar (tests/f64_bar.py) admits another summation
    order of the oracle's own reductions and rejects a plausible mistake of a
    double/double path: one input rounded through float (here b).  The chunk
    length changes only the order of dots and norms: the default restores
    bit-identical results and the products do not depend on it."""
    import f64_bar as F
    A, x, b = _dense(20, 20, 0.5, 17)
    op = O.Op.dense(A)
    rng = np.random.default_rng(3)
    bd = b.astype(np.float64) + 1e-3 * rng.standard_normal(b.size)  # not float-valued
    opts = default_options(maxit=60, ls_exact=1)
    Uo, io, _ = O.solve_f64(op, bd, (20, 20, 1), opts)
    bar = F.bound(op, bd, (20, 20, 1), opts, Uo)
    O.set_sum_chunk(100)
    try:
        Uc, ic, _ = O.solve_f64(op, bd, (20, 20, 1), opts)
        y = op.gemv_f64(bd, trans=True)
    finally:
        O.set_sum_chunk(4096)
    assert F.rel(Uc, Uo) <= bar and ic.total_iters == io.total_iters
    np.testing.assert_array_equal(y, op.gemv_f64(bd, trans=True))
    U2, _, _ = O.solve_f64(op, bd, (20, 20, 1), opts)
    assert F.rel(U2, Uo) == 0.0
    Uf, _, _ = O.solve_f64(op, bd.astype(np.float32).astype(np.float64), (20, 20, 1), opts)
    assert F.rel(Uf, Uo) > 10 * bar, (F.rel(Uf, Uo), bar)


# ---------------------------------------------------------------------------
# step size, line search and multipliers, pinned by the mathematics of the
# augmented Lagrangian (no reference formula retyped):
#   L(u, w; s, dl) = sum|w| - s.(Du - w) + beta/2 |Du - w|^2
#                    - dl.(Au - b) + mu/2 |Au - b|^2
# The solver's d is grad_u L / beta; its first step of every outer iteration
# is the exact minimiser of L along d; later steps are Barzilai-Borwein
# (s's / s'y on consecutive iterates and gradients); accepted steps satisfy
# the nonmonotone Armijo rule L_new <= C - c alpha tau beta |d|^2 and the
# rejected alpha / gamma did not; multipliers follow s -= beta (Du - w),
# dl -= mu (Au - b).
# ---------------------------------------------------------------------------
class _ALM:
    def __init__(self, A, bs, q, p, mu, beta):
        self.A = A.astype(np.float64)
        self.bs, self.q, self.p, self.mu, self.beta = bs, q, p, mu, beta

    def D(self, U):
        U2 = U.reshape(self.p, self.q)
        return np.concatenate([(np.roll(U2, -1, 1) - U2).ravel(),
                               (np.roll(U2, -1, 0) - U2).ravel()])

    def Dt(self, X):
        n = self.p * self.q
        X2, Y2 = X[:n].reshape(self.p, self.q), X[n:].reshape(self.p, self.q)
        return ((np.roll(X2, 1, 1) - X2) + (np.roll(Y2, 1, 0) - Y2)).ravel()

    def W(self, U, s):
        v = self.D(U) - s / self.beta
        return np.sign(v) * np.maximum(np.abs(v) - 1.0 / self.beta, 0.0)

    def L(self, U, W, s, dl):
        V, r = self.D(U) - W, self.A @ U - self.bs
        return (np.abs(W).sum() - s @ V + self.beta / 2 * V @ V
                - dl @ r + self.mu / 2 * r @ r)

    def d(self, U, W, s, dl):
        V, r = self.D(U) - W, self.A @ U - self.bs
        return (self.Dt(-s + self.beta * V) + self.A.T @ (-dl + self.mu * r)) / self.beta

    def line_min(self, U, W, s, dl, d):
        """exact minimiser of the quadratic t -> L(U - t d) from three values"""
        h = 1e-3 * np.linalg.norm(U) / max(np.linalg.norm(d), 1e-300)
        f0, f1, f2 = (self.L(U - t * d, W, s, dl) for t in (0.0, h, 2 * h))
        return h * (3 * f0 - 4 * f1 + f2) / (2 * (f0 - 2 * f1 + f2))


def _alm_setup(seed, c=1e-5):
    q, p = 12, 10
    A, x, b = _dense(q, p, 0.5, seed)
    s = _scaleb(b)
    bs = s * b.astype(np.float64)
    opts = default_options(scal_double=1, ls_exact=1, c=c)
    M = _ALM(A, bs, q, p, opts.mu, opts.beta)
    return A, b, s, opts, M, (q, p, 1)


def _run_f64(A, b, dims, opts, maxit, s, trace=0):
    opts.maxit = maxit
    U, info, tr = O.solve_f64(O.Op.dense(A), b.astype(np.float64), dims, opts,
                              trace_cap=trace)
    return U * s, info, tr


# c = 0.6: at the exact line minimum of a quadratic the decrease is half the
# linear slope, so alpha = 1 fails the rule and alpha = gamma passes
@pytest.mark.parametrize("seed,c", [(21, 1e-5), (22, 1e-5), (23, 1e-5), (21, 0.6), (23, 0.6)])
def test_first_step_bb_step_and_armijo(seed, c):
    A, b, s, opts, M, dims = _alm_setup(seed, c)
    n = dims[0] * dims[1]
    z2, zm = np.zeros(2 * n), np.zeros(A.shape[0])
    U0 = M.A.T @ M.bs
    W0 = M.W(U0, z2)
    d0 = M.d(U0, W0, z2, zm)
    U1, _, tr = _run_f64(A, b, dims, opts, 1, s, trace=1)
    t1 = tr[0]
    assert not t1.fallback
    if c > 0.5:
        assert t1.backtracks == 1
    # first step: exact minimiser along d
    assert t1.tau == pytest.approx(M.line_min(U0, W0, z2, zm, d0), rel=1e-7)
    # the accepted step is U0 - alpha tau d with alpha = gamma^k
    assert t1.alpha == pytest.approx(opts.gamma ** t1.backtracks, rel=1e-12)
    np.testing.assert_allclose(U1, U0 - t1.alpha * t1.tau * d0, rtol=1e-9,
                               atol=1e-9 * np.abs(U0).max())
    # nonmonotone Armijo acceptance (C = L at the start)
    C0 = M.L(U0, W0, z2, zm)
    slope = opts.c * t1.tau * opts.beta * (d0 @ d0)

    def ok(al):
        return abs(M.L(U0 - al * t1.tau * d0, W0, z2, zm)) <= abs(C0 - al * slope)
    assert ok(t1.alpha)
    if t1.backtracks:
        assert not ok(t1.alpha / opts.gamma)
    # f after the w-update is L at the new (u, w)
    W1 = M.W(U1, z2)
    assert t1.f == pytest.approx(M.L(U1, W1, z2, zm), rel=1e-9)
    # second step: Barzilai-Borwein s's / s'y on the gradients d
    d1 = M.d(U1, W1, z2, zm)
    U2, _, tr2 = _run_f64(A, b, dims, opts, 2, s, trace=2)
    sv, yv = U1 - U0, d1 - d0
    assert tr2[1].tau == pytest.approx(abs(sv @ sv / (sv @ yv)), rel=1e-7)
    if not tr2[1].fallback:
        np.testing.assert_allclose(U2, U1 - tr2[1].alpha * tr2[1].tau * d1, rtol=1e-9,
                                   atol=1e-9 * np.abs(U1).max())


@pytest.mark.parametrize("seed", [24, 25])
def test_multiplier_update_sets_next_outer_gradient(seed):
    """After the first outer iteration s = -beta (DU - W), dl = -mu (AU - b);
    the next outer iteration's first step is the exact minimiser of L with
    those multipliers along d = grad_u L / beta."""
    A, b, s, opts, M, dims = _alm_setup(seed)
    _, info, tr = _run_f64(A, b, dims, opts, 80, s, trace=80)
    k = next(i for i, t in enumerate(tr) if t.outer == 1)
    assert k >= 2
    Uk, _, _ = _run_f64(A, b, dims, opts, k, s)
    n = dims[0] * dims[1]
    Wk = M.W(Uk, np.zeros(2 * n))
    sg = -M.beta * (M.D(Uk) - Wk)
    dl = -M.mu * (M.A @ Uk - M.bs)
    d = M.d(Uk, Wk, sg, dl)
    assert tr[k].tau == pytest.approx(M.line_min(Uk, Wk, sg, dl, d), rel=1e-6)
    assert tr[k - 1].outer == 0
    # the step is taken along that d, and the w-update shrinks DU - s / beta:
    # f after it is L(U, W; s, dl)
    Uk1, _, _ = _run_f64(A, b, dims, opts, k + 1, s)
    t = tr[k]
    if not t.fallback:
        np.testing.assert_allclose(Uk1, Uk - t.alpha * t.tau * d, rtol=1e-8,
                                   atol=1e-8 * np.abs(Uk).max())
    Wk1 = M.W(Uk1, sg)
    assert t.f == pytest.approx(M.L(Uk1, Wk1, sg, dl), rel=1e-8)


def test_nonneg_clamps_every_iterate():
    A, x, b = _dense(16, 12, 0.4, 26)
    opts = default_options(maxit=40, nonneg=1, scal_double=1)
    U, info, _ = O.solve(O.Op.dense(A), b, (16, 12, 1), opts)
    assert info.total_iters > 0 and U.min() >= 0.0
    opts.nonneg = 0
    U2, _, _ = O.solve(O.Op.dense(A), b, (16, 12, 1), opts)
    assert U2.min() < 0.0  # the clamp is doing something on this problem


@pytest.mark.parametrize("seed", [27, 28])
def test_quadratic_model_line_search_matches_rounds(seed):
    """With W fixed during the line search, f(alpha) is an exact quadratic in
    alpha, so deciding all backtracking rounds from its coefficients
    (ls_exact = 0) must take the same steps as evaluating every round
    (ls_exact = 1); in FP64 the trajectories agree to rounding."""
    A, x, b = _dense(16, 16, 0.4, seed)
    op = O.Op.dense(A)
    res = []
    for ls in (1, 0):
        opts = default_options(maxit=60, scal_double=1, ls_exact=ls, c=0.6)
        U, info, tr = O.solve_f64(op, b.astype(np.float64), (16, 16, 1), opts, trace_cap=60)
        res.append((U, info, tr))
    (U1, i1, t1), (U0, i0, t0) = res
    assert sum(t.backtracks for t in t1) > 5
    assert [t.backtracks for t in t1] == [t.backtracks for t in t0]
    assert i1.total_iters == i0.total_iters
    assert np.linalg.norm(U1 - U0) / np.linalg.norm(U1) < 1e-10


@pytest.mark.parametrize("seed", [23, 24])
def test_five_strike_fallback(seed):
    """c = 0.97 on these problems: the first step of the second outer
    iteration fails the rule five times (the rule needs alpha <= 2 (1 - c) =
    0.06 at the exact line minimum; gamma^5 = 0.078).  The fallback takes the
    full first-step tau along d, and the nonmonotone weight gam is multiplied
    by rate_gam in that step's C average (Q restarts at 1, C at L)."""
    A, b, s, opts, M, dims = _alm_setup(seed, c=0.97)
    _, _, tr = _run_f64(A, b, dims, opts, 80, s, trace=80)
    k = next(i for i, t in enumerate(tr) if t.outer == 1)
    t = tr[k]
    assert t.fallback == 1
    Uk, _, _ = _run_f64(A, b, dims, opts, k, s)
    n = dims[0] * dims[1]
    Wk = M.W(Uk, np.zeros(2 * n))
    sg = -M.beta * (M.D(Uk) - Wk)
    dl = -M.mu * (M.A @ Uk - M.bs)
    d = M.d(Uk, Wk, sg, dl)
    tau = M.line_min(Uk, Wk, sg, dl, d)
    assert t.tau == pytest.approx(tau, rel=1e-6)
    Uk1, _, _ = _run_f64(A, b, dims, opts, k + 1, s)
    np.testing.assert_allclose(Uk1, Uk - tau * d, rtol=1e-8, atol=1e-8 * np.abs(Uk).max())
    C0 = M.L(Uk, Wk, sg, dl)
    g = opts.gam * opts.rate_gam
    assert t.C == pytest.approx((g * C0 + t.f) / (g + 1), rel=1e-8)


@pytest.mark.parametrize("n,m,smb", [(4096, 1229, 5), (1024, 700, 1), (512, 300, 0)])
def test_dense_get_g_blocking(n, m, smb):
    """get_g_mult's dense product (2d/tval3_cpu_2d.cpp:361-394): row blocks of
    max(submat_mb MiB / 4 / n, 1) rows; each block's A_blk^T Au_blk is an FP64
    sum rounded to float, the blocks are added in float in order.  Checked
    against that formula in numpy, and against the unblocked product, which
    it must differ from in the last bits when there are several blocks."""
    rng = np.random.default_rng(n + m)
    A = (rng.standard_normal((m, n)) / np.sqrt(m)).astype(np.float32)
    U = rng.standard_normal(n).astype(np.float32)
    op = O.Op.dense(A)
    Au, g = op.AtA_blocked(U, smb)
    A64 = A.astype(np.float64)
    Au_ref = (A64 @ U.astype(np.float64)).astype(np.float32)
    np.testing.assert_array_equal(Au, Au_ref)
    r = max((max(smb, 0) * 1024 * 1024 // 4) // n, 1)
    gref = np.zeros(n, np.float32)
    for y0 in range(0, m, r):
        blk = (A64[y0:y0 + r].T @ Au_ref[y0:y0 + r].astype(np.float64)).astype(np.float32)
        gref = (gref + blk).astype(np.float32)
    assert np.max(np.abs(g.astype(np.float64) - gref)) <= 2e-7 * np.abs(gref).max()
    nblk = (m + r - 1) // r
    plain = op.gemv(Au, trans=True)
    if nblk > 2:
        assert np.any(plain != g)


def _layered_case(L, q, p_layer, frac, seed, sparse):
    """L layers of a q x p_layer image stacked along y, one A for all of them;
    returns the layered operator, the explicit block-diagonal operator
    I_L (x) A, b and the image dims."""
    import scipy.sparse as sp
    N = q * p_layer
    M = int(round(frac * N))
    A = synth.gaussian_A(M, N, seed)
    if sparse:
        A[np.abs(A) < 0.8 * np.abs(A).mean()] = 0
    x = np.concatenate([synth.ellipse_phantom(p_layer, q, 4, seed + l).ravel() for l in range(L)])
    Abig = np.kron(np.eye(L, dtype=np.float32), A).astype(np.float32)
    b = synth.add_noise(Abig.astype(np.float64) @ x, 0.01, seed).astype(np.float32)
    if sparse:
        S, T = sp.csr_matrix(A), sp.csc_matrix(A)
        Sb, Tb = sp.csr_matrix(Abig), sp.csc_matrix(Abig)
        op = O.Op.sparse(M, N, S.indptr, S.indices, S.data, T.indptr, T.indices, T.data)
        big = O.Op.sparse(L * M, L * N, Sb.indptr, Sb.indices, Sb.data, Tb.indptr, Tb.indices, Tb.data)
    else:
        op, big = O.Op.dense(A), O.Op.dense(Abig)
    return op, big, b, (q, L * p_layer, 1)


@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("sd", [0, 1])
def test_layers_equal_block_diagonal_operator(sparse, sd):
    """The 2-D multi-layer mode (layers = b.len / A.dim_y; 2d/mat_mult.h:22-35,
    68-90 multiply every layer by A; 2d:365 and 390-393 take two plain
    products in get_g) is the single-layer solve with the block-diagonal
    operator I_L (x) A.  With one row block (submat_mb large) the blocked get_g
    of that solve is the plain pair of products too, and the zero blocks add
    exact zeros in order, so the two solves agree bit for bit."""
    op, big, b, dims = _layered_case(3, 12, 10, 0.4, 7, sparse)
    opts = default_options(maxit=60, scal_double=sd, submat_mb=4096)
    U1, i1, t1 = O.solve(op, b, dims, opts, trace_cap=100)
    U2, i2, t2 = O.solve(big, b, dims, opts, trace_cap=100)
    np.testing.assert_array_equal(U1, U2)
    assert i1.total_iters == i2.total_iters and i1.outer_iters == i2.outer_iters
    assert [t.f for t in t1] == [t.f for t in t2]


def test_layers_take_plain_products():
    """Several layers never take the blocked get_g product (2d:365): 1 MiB row
    blocks change nothing there, while they do change the single-layer solve
    of the block-diagonal operator (720 rows in blocks of 145)."""
    op, big, b, dims = _layered_case(2, 30, 30, 0.4, 5, False)
    o1, o2 = default_options(maxit=20, submat_mb=1), default_options(maxit=20, submat_mb=4096)
    np.testing.assert_array_equal(O.solve(op, b, dims, o1)[0], O.solve(op, b, dims, o2)[0])
    assert np.any(O.solve(big, b, dims, o1)[0] != O.solve(big, b, dims, o2)[0])


def test_layers_argument_checks():
    op, _, b, (q, p, r) = _layered_case(2, 8, 6, 0.5, 3, False)
    opts = default_options(maxit=5)
    with pytest.raises(ValueError):       # b is one layer short
        O.solve(op, b[: op.m], (q, p, r), opts)
    with pytest.raises(ValueError):       # U is not a whole number of layers
        O.solve(op, b, (q, p + 1, r), opts)
    geo = synth.ring_geometry(16, 16)
    gop = O.Op.geometry(geo)
    with pytest.raises(ValueError):       # no layers for the projector
        O.solve(gop, np.zeros(2 * gop.m, np.float32), (16, 32, 1), opts)


@pytest.mark.parametrize("skip,extra", [(0, 0), (3, 7)])
def test_multi_mp_rows_are_per_position_operators(skip, extra):
    """numMPs > 1 (3d/mat_mult.h:183-202): row i * num_paths + use_path - 1 is
    the walk of the pair at measurement position i, so the multi-position
    operator is the per-position single-position operators stacked, row for
    row (walks and products bit for bit); rows no pair maps to are 0."""
    K, grid, P = 24, 32, 3
    geo = synth.multi_mp_ring_geometry(K, grid, P, skip_every=skip, extra_rows=extra, seed=5)
    op = O.Op.geometry(geo)
    ptr, lin = op.walk()
    x = synth.disk_phantom(grid, seed=2)
    y = op.gemv(x)
    npaths = geo.num_paths
    for i in range(P):
        sl = slice(i * K, (i + 1) * K)
        one = synth.GeometryArrays(1, geo.use_path, geo.xe[sl], geo.ye[sl], geo.ze[sl],
                                   geo.xr[sl], geo.yr[sl], geo.zr[sl], geo.rv, npaths)
        op1 = O.Op.geometry(one)
        p1, l1 = op1.walk()
        base = i * npaths
        for r in range(npaths):
            np.testing.assert_array_equal(lin[ptr[base + r]:ptr[base + r + 1]], l1[p1[r]:p1[r + 1]])
        np.testing.assert_array_equal(y[base:base + npaths], op1.gemv(x))
        used = np.zeros(npaths, bool)
        used[geo.use_path[geo.use_path > 0] - 1] = True
        assert np.count_nonzero(~used) == extra and np.all(y[base:base + npaths][~used] == 0)
    # the transpose gathers the same rays
    yb = synth.add_noise(y, 0.1, 1)
    At = op.gemv(yb, trans=True)
    ref = sum(O.Op.geometry(synth.GeometryArrays(
        1, geo.use_path, geo.xe[i * K:(i + 1) * K], geo.ye[i * K:(i + 1) * K], geo.ze[i * K:(i + 1) * K],
        geo.xr[i * K:(i + 1) * K], geo.yr[i * K:(i + 1) * K], geo.zr[i * K:(i + 1) * K], geo.rv,
        npaths)).gemv(yb[i * npaths:(i + 1) * npaths], trans=True).astype(np.float64) for i in range(P))
    np.testing.assert_allclose(At, ref, rtol=1e-6, atol=1e-5)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_line_search_modes_agree_on_ring256(seed):
    """The 256x256 ring workload (256 transducers, maxit 200) in FP32: the
    quadratic-model line search (ls_exact = 0, the engine default) and the
    reference's round-by-round backtracking (ls_exact = 1) give the same
    reconstruction bit for bit and the same trajectory."""
    geo = synth.ring_geometry(256, 256)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(256, seed=0)
    b = synth.add_noise(op.gemv(x), 0.005, seed=seed)
    res = []
    for ls in (0, 1):
        opts = default_options(maxit=200, ls_exact=ls)
        U, info, tr = O.solve(op, b, (256, 256, 1), opts, trace_cap=200)
        res.append((U, info, [(t.backtracks, t.fallback) for t in tr]))
    (U0, i0, t0), (U1, i1, t1) = res
    assert t0 == t1 and i0.total_iters == i1.total_iters
    assert sum(fb for _, fb in t0) > 0
    np.testing.assert_array_equal(U0.view(np.uint32), U1.view(np.uint32))
