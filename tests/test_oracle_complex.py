"""The oracle's complex instantiations (the reference's mat<complex<float>> /
mat<complex<double>> overloads), pinned without retyping their formulas:

* products against numpy (dense and sparse complex A, A^H by the adjoint
  identity), the projector per component against the real projector;
* the complex shrinkage against its closed form (soft threshold of the
  complex magnitude per direction, and the reference's 3-D Wz, which takes
  the y direction's ubar);
* an FP64 numpy model of the complex augmented Lagrangian: the first step is
  the exact line minimiser along d = grad L / beta (A^H, the complex
  shrinkage and the real parts of the dots all enter), the second the
  Barzilai-Borwein |s^H s / s^H y| (s^H y is nearly real here: the smooth
  part of the Hessian is Hermitian);
* symmetries: multiplying b by i multiplies U by i bit for bit; a general
  phase e^{i theta} agrees to FP64 rounding; real data reproduces the real
  solve with zero imaginary parts;
* nonneg and isreal leave no imaginary part; isotropic TV and operators
  without complex entries are refused.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from tval3_amd import synth
from tval3_amd.abi import default_options
import tv3_oracle as O


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def _crand(rng, shape, t=np.complex128):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(t)


@pytest.mark.parametrize("t,tol", [(np.complex64, 1e-6), (np.complex128, 1e-14)])
def test_complex_products_against_numpy(t, tol):
    rng = np.random.default_rng(1)
    m, n = 70, 130
    A = _crand(rng, (m, n), t)
    x, y = _crand(rng, n, t), _crand(rng, m, t)
    Ad = A.astype(np.complex128)
    op = O.Op.dense_c(A)
    assert _rel(op.gemv_c(x), Ad @ x) < tol
    assert _rel(op.gemv_c(y, trans=True), Ad.conj().T @ y) < tol
    # adjoint identity <A x, y> = <x, A^H y>
    lhs = np.vdot(op.gemv_c(x).astype(np.complex128), y)
    rhs = np.vdot(x.astype(np.complex128), op.gemv_c(y, trans=True))
    assert abs(lhs - rhs) < 10 * tol * abs(lhs)
    As = A * (rng.random((m, n)) < 0.15)
    S, T = sp.csr_matrix(As), sp.csc_matrix(As)
    sop = O.Op.sparse_c(m, n, S.indptr, S.indices, S.data, T.indptr, T.indices, T.data)
    Asd = As.astype(np.complex128)
    assert _rel(sop.gemv_c(x), Asd @ x) < tol
    assert _rel(sop.gemv_c(y, trans=True), Asd.conj().T @ y) < tol


def test_complex_projector_is_per_component():
    rng = np.random.default_rng(2)
    for geo in (synth.ring_geometry(32, 32), synth.hemisphere_geometry(24, 10)):
        op = O.Op.geometry(geo)
        x = _crand(rng, geo.n, np.complex64)
        y = _crand(rng, geo.m, np.complex64)
        ax, ay = op.gemv_c(x), op.gemv_c(y, trans=True)
        np.testing.assert_array_equal(ax.real, op.gemv(x.real))
        np.testing.assert_array_equal(ax.imag, op.gemv(x.imag))
        np.testing.assert_array_equal(ay.real, op.gemv(y.real, trans=True))
        np.testing.assert_array_equal(ay.imag, op.gemv(y.imag, trans=True))


@pytest.mark.parametrize("dims", [(8, 6, 1), (5, 4, 3)])
@pytest.mark.parametrize("wz_fix", [0, 1])
def test_complex_shrink_closed_form(dims, wz_fix):
    """W_dir = max(|ubar| - 1/beta, 0) ubar / |ubar| with ubar = U - sigma /
    beta, lam1 = sum of the thresholded magnitudes; the reference's 3-D Wz
    scales the y direction's ubar (wz_fix = 0)."""
    rng = np.random.default_rng(3 + wz_fix)
    n = int(np.prod(dims))
    z3 = dims[2] > 1
    beta = 0.7
    U = [_crand(rng, n, np.complex64) for _ in range(3)]
    sg = [_crand(rng, n, np.complex64) * 2 for _ in range(3)]
    Wx, Wy, Wz, lam1 = O.shrink_c(U[0], U[1], U[2] if z3 else None, sg[0], sg[1],
                                  sg[2] if z3 else None, dims, beta, wz_fix)
    bars = [(u.astype(np.complex128) - s.astype(np.complex128) / beta) for u, s in zip(U, sg)]
    mags = [np.abs(b) for b in bars]
    ws = [np.maximum(mg - 1 / beta, 0) for mg in mags]
    np.testing.assert_allclose(Wx, ws[0] * bars[0] / mags[0], rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(Wy, ws[1] * bars[1] / mags[1], rtol=2e-6, atol=1e-6)
    tot = ws[0].sum() + ws[1].sum()
    if z3:
        src = bars[2] if wz_fix else bars[1]
        np.testing.assert_allclose(Wz, ws[2] * src / mags[2], rtol=2e-6, atol=1e-6)
        tot += ws[2].sum()
    assert lam1 == pytest.approx(tot, rel=1e-5)
    assert (ws[0] == 0).any() and (ws[0] > 0).any()  # both branches taken


# ---- the complex augmented Lagrangian --------------------------------------
class _CALM:
    """L(u, w; s, dl) = sum_dir |w| - Re s^H (Du - w) + beta/2 |Du - w|^2
    - Re dl^H (Au - b) + mu/2 |Au - b|^2 with complex u, w, s, dl."""

    def __init__(self, A, bs, q, p, mu, beta):
        self.A, self.bs, self.q, self.p, self.mu, self.beta = A, bs, q, p, mu, beta

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
        a = np.abs(v)
        return np.where(a > 0, np.maximum(a - 1 / self.beta, 0) * v / np.where(a > 0, a, 1), 0)

    def L(self, U, W, s, dl):
        V, r = self.D(U) - W, self.A @ U - self.bs
        return (np.abs(W).sum() - np.vdot(s, V).real + self.beta / 2 * np.vdot(V, V).real
                - np.vdot(dl, r).real + self.mu / 2 * np.vdot(r, r).real)

    def d(self, U, W, s, dl):
        V, r = self.D(U) - W, self.A @ U - self.bs
        return (self.Dt(-s + self.beta * V) + self.A.conj().T @ (-dl + self.mu * r)) / self.beta

    def line_min(self, U, W, s, dl, d):
        h = 1e-3 * np.linalg.norm(U) / max(np.linalg.norm(d), 1e-300)
        f0, f1, f2 = (self.L(U - t * d, W, s, dl) for t in (0.0, h, 2 * h))
        return h * (3 * f0 - 4 * f1 + f2) / (2 * (f0 - 2 * f1 + f2))


def _scaleb_c(b):
    a = np.abs(b)
    bmin, bmax = b[np.argmin(a)], b[np.argmax(a)]
    dif = abs(bmax - bmin)
    return 0.5 / dif if dif < 0.5 else (1.5 / dif if dif > 1.5 else 1.0)


@pytest.mark.parametrize("seed", [31, 32, 33])
def test_complex_first_step_and_bb_step(seed):
    rng = np.random.default_rng(seed)
    q, p = 12, 10
    n, m = q * p, 60
    A = _crand(rng, (m, n)) / np.sqrt(m)
    x = _crand(rng, n)
    b = A @ x + 0.01 * _crand(rng, m)
    opts = default_options(ls_exact=1, isreal=0)
    s = _scaleb_c(b)
    M = _CALM(A, s * b, q, p, opts.mu, opts.beta)
    z2, zm = np.zeros(2 * n, complex), np.zeros(m, complex)
    U0 = A.conj().T @ M.bs
    W0 = M.W(U0, z2)
    d0 = M.d(U0, W0, z2, zm)
    op = O.Op.dense_c(A)
    opts.maxit = 1
    U1, _, tr = O.solve_c(op, b, (q, p, 1), opts, trace_cap=1)
    U1 = U1 * s
    t1 = tr[0]
    assert not t1.fallback
    assert t1.tau == pytest.approx(M.line_min(U0, W0, z2, zm, d0), rel=1e-7)
    np.testing.assert_allclose(U1, U0 - t1.alpha * t1.tau * d0, rtol=1e-9, atol=1e-9 * np.abs(U0).max())
    W1 = M.W(U1, z2)
    assert t1.f == pytest.approx(M.L(U1, W1, z2, zm), rel=1e-9)
    d1 = M.d(U1, W1, z2, zm)
    opts.maxit = 2
    _, _, tr2 = O.solve_c(op, b, (q, p, 1), opts, trace_cap=2)
    sv, yv = U1 - U0, d1 - d0
    assert tr2[1].tau == pytest.approx(abs(np.vdot(sv, sv) / np.vdot(sv, yv)), rel=1e-7)


# ---- symmetries -------------------------------------------------------------
def _cases():
    rng = np.random.default_rng(5)
    Ar = synth.gaussian_A(150, 256, 6)
    A = (Ar + 1j * synth.gaussian_A(150, 256, 7)).astype(np.complex64)
    xx = synth.ellipse_phantom(16, 16, 4, 6).ravel()
    b = (A.astype(np.complex128) @ xx + 0.01 * _crand(rng, 150)).astype(np.complex64)
    yield "dense", O.Op.dense_c(A), b, (16, 16, 1)
    As = A * (rng.random(A.shape) < 0.3)
    S, T = sp.csr_matrix(As), sp.csc_matrix(As)
    yield "sparse", O.Op.sparse_c(150, 256, S.indptr, S.indices, S.data, T.indptr, T.indices,
                                  T.data), b, (16, 16, 1)
    geo = synth.ring_geometry(24, 24)
    op = O.Op.geometry(geo)
    x = synth.disk_phantom(24, seed=1)
    bg = (op.gemv(x) + 1j * op.gemv(synth.disk_phantom(24, seed=2))).astype(np.complex64)
    yield "ring", op, bg, (24, 24, 1)
    geo3 = synth.hemisphere_geometry(30, 10)
    op3 = O.Op.geometry(geo3)
    x3 = synth.ellipsoid_phantom(10, 3, seed=1)
    b3 = (op3.gemv(x3) * (1 + 0.5j)).astype(np.complex64)
    yield "hemisphere", op3, b3, (10, 10, 10)


@pytest.mark.parametrize("exact", [0, 1])
def test_multiplication_by_i_is_exact(exact):
    for name, op, b, dims in _cases():
        opts = default_options(maxit=30, ls_exact=1, isreal=0, dt_exact=exact)
        U1, i1, t1 = O.solve_c(op, b, dims, opts, trace_cap=30)
        U2, i2, t2 = O.solve_c(op, (1j * b).astype(np.complex64), dims, opts, trace_cap=30)
        np.testing.assert_array_equal(U2, (1j * U1).astype(np.complex64), err_msg=name)
        assert [t.tau for t in t1] == [t.tau for t in t2], name


def test_phase_covariance_and_real_data():
    for name, op, b, dims in _cases():
        opts = default_options(maxit=30, ls_exact=1, isreal=0)
        b2 = b.astype(np.complex128)
        U1, _, _ = O.solve_c(op, b2, dims, opts)
        U2, _, _ = O.solve_c(op, np.exp(0.7j) * b2, dims, opts)
        assert _rel(U2, np.exp(0.7j) * U1) < 1e-9, name
    # real data: the imaginary part stays 0 and the real part is the real
    # solve (FP64: to rounding; the complex magnitude in the shrinkage
    # rounds differently from the real absolute value)
    Ar = synth.gaussian_A(200, 400, 8)
    xx = synth.ellipse_phantom(20, 20, 4, 8).ravel()
    b = synth.add_noise(Ar @ xx, 0.01, 8)
    opts = default_options(maxit=40, ls_exact=1, isreal=0)
    Ur, _, _ = O.solve_f64(O.Op.dense(Ar), b.astype(np.float64), (20, 20, 1), opts)
    Uc, _, _ = O.solve_c(O.Op.dense_c(Ar), b.astype(np.complex128), (20, 20, 1), opts)
    assert np.abs(Uc.imag).max() == 0.0
    assert _rel(Uc.real, Ur) < 1e-10


def test_nonneg_isreal_and_refusals():
    name, op, b, dims = next(_cases())
    opts = default_options(maxit=20, nonneg=1, isreal=0)
    U, _, _ = O.solve_c(op, b, dims, opts)
    assert np.abs(U.imag).max() == 0.0 and U.real.min() >= 0.0
    opts = default_options(maxit=20, isreal=1)
    U, _, _ = O.solve_c(op, b, dims, opts)
    assert np.abs(U.imag).max() == 0.0
    opts.isreal = 0
    U, _, _ = O.solve_c(op, b, dims, opts)
    assert np.abs(U.imag).max() > 0.0
    with pytest.raises(ValueError, match="isotropic"):
        O.solve_c(op, b, dims, default_options(maxit=5, tv_iso=1))
    with pytest.raises(ValueError, match="complex entries"):
        O.solve_c(O.Op.dense(synth.gaussian_A(150, 256, 1)), b, dims, default_options(maxit=5))
