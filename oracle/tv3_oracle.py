"""ctypes wrapper of the CPU oracle (oracle/libtv3oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, always as the checker or the CPU reference timing,
never as part of the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtv3oracle.so")
_lib = None

_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)
_LP = C.POINTER(C.c_longlong)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.tv3o_op_dense.restype = C.c_void_p
        L.tv3o_op_dense.argtypes = [C.c_int, C.c_int, _FP]
        L.tv3o_op_sparse.restype = C.c_void_p
        L.tv3o_op_sparse.argtypes = [C.c_int, C.c_int, C.c_int, _IP, _IP, _FP,
                                     _IP, _IP, _FP]
        L.tv3o_op_geometry.restype = C.c_void_p
        L.tv3o_op_geometry.argtypes = [C.c_void_p]
        L.tv3o_op_free.argtypes = [C.c_void_p]
        L.tv3o_op_nnz.restype = C.c_longlong
        L.tv3o_op_nnz.argtypes = [C.c_void_p]
        L.tv3o_op_walk.argtypes = [C.c_void_p, _LP, _IP]
        L.tv3o_gemv.argtypes = [C.c_void_p, C.c_int, _FP, _FP]
        L.tv3o_D.argtypes = [C.c_int] * 3 + [_FP] * 4
        L.tv3o_Dt.argtypes = [C.c_int] * 4 + [_FP] * 4
        L.tv3o_shrink.restype = C.c_double
        L.tv3o_shrink.argtypes = [C.c_int] * 3 + [C.c_double, C.c_int, C.c_int] + [_FP] * 9
        L.tv3o_dot.restype = C.c_double
        L.tv3o_dot.argtypes = [C.c_longlong, _FP, _FP]
        L.tv3o_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                 C.c_int, _FP, _FP, _FP, C.c_void_p,
                                 C.c_void_p, C.c_int]
        _DP = C.POINTER(C.c_double)
        L.tv3o_solve_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                     C.c_int, _DP, _DP, _DP, C.c_void_p,
                                     C.c_void_p, C.c_int]
        L.tv3o_gemv_f64.argtypes = [C.c_void_p, C.c_int, _DP, _DP]
        L.tv3o_AtA_blocked.argtypes = [C.c_void_p, _FP, C.c_int, _FP, _FP]
        L.tv3o_set_threads.argtypes = [C.c_int]
        L.tv3o_set_sum_chunk.argtypes = [C.c_longlong]
        L.tv3o_op_dense_c.restype = C.c_void_p
        L.tv3o_op_dense_c.argtypes = [C.c_int, C.c_int, _DP]
        L.tv3o_op_sparse_c.restype = C.c_void_p
        L.tv3o_op_sparse_c.argtypes = [C.c_int, C.c_int, C.c_int, _IP, _IP, _DP,
                                       _IP, _IP, _DP]
        for nm, P_ in (("c64", _FP), ("c128", _DP)):
            getattr(L, "tv3o_solve_" + nm).argtypes = [
                C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, P_, P_, P_,
                C.c_void_p, C.c_void_p, C.c_int]
            getattr(L, "tv3o_gemv_" + nm).argtypes = [C.c_void_p, C.c_int, P_, P_]
        L.tv3o_shrink_c64.restype = C.c_double
        L.tv3o_shrink_c64.argtypes = [C.c_int] * 3 + [C.c_double, C.c_int] + [_FP] * 9
        L.tv3o_get_threads.restype = C.c_int
        L.tv3o_plant.argtypes = [C.c_int]
        L.tv3o_counts.argtypes = [_LP, _LP]
        _lib = L
    return _lib


def _f(a):
    return None if a is None else a.ctypes.data_as(_FP)


def _i(a):
    return a.ctypes.data_as(_IP)


def _c32(a):
    return np.ascontiguousarray(a, np.float32)


def set_threads(t):
    lib().tv3o_set_threads(int(t))


def get_threads():
    return lib().tv3o_get_threads()


def set_sum_chunk(c):
    """Chunk length of the FP64 reductions (default 4096): only the
    summation order of dots and norms changes."""
    lib().tv3o_set_sum_chunk(int(c))


def plant(k):
    """Plants oracle mistake k (0: none); for the tests that show each one
    fails a pinning check."""
    lib().tv3o_plant(int(k))


def reset_counts():
    lib().tv3o_reset_counts()


def counts():
    """(A x, A^T x) products the oracle's solver issued since reset_counts()."""
    f, b = C.c_longlong(0), C.c_longlong(0)
    lib().tv3o_counts(C.byref(f), C.byref(b))
    return f.value, b.value


class Op:
    """A measurement operator held by the oracle (keeps its arrays alive)."""

    def __init__(self, handle, m, n, keep):
        self.h = handle
        self.m, self.n = m, n
        self._keep = keep

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.tv3o_op_free(self.h)
            self.h = None

    @classmethod
    def dense(cls, A):
        A = _c32(A)
        m, n = A.shape
        return cls(lib().tv3o_op_dense(m, n, _f(A)), m, n, [A])

    @classmethod
    def sparse(cls, m, n, csr_ptr, csr_ind, csr_val, csc_ptr, csc_ind, csc_val):
        arrs = [np.ascontiguousarray(a, t) for a, t in
                ((csr_ptr, np.int32), (csr_ind, np.int32), (csr_val, np.float32),
                 (csc_ptr, np.int32), (csc_ind, np.int32), (csc_val, np.float32))]
        h = lib().tv3o_op_sparse(m, n, int(arrs[1].size), _i(arrs[0]),
                                 _i(arrs[1]), _f(arrs[2]), _i(arrs[3]),
                                 _i(arrs[4]), _f(arrs[5]))
        return cls(h, m, n, arrs)

    @classmethod
    def dense_c(cls, A):
        """Complex dense A (the complex overloads): entries kept as complex128
        (complex64 entries widen exactly)."""
        A = np.ascontiguousarray(A, np.complex128)
        m, n = A.shape
        return cls(lib().tv3o_op_dense_c(m, n, A.view(np.float64).ctypes.data_as(
            C.POINTER(C.c_double))), m, n, [A])

    @classmethod
    def sparse_c(cls, m, n, csr_ptr, csr_ind, csr_val, csc_ptr, csc_ind, csc_val):
        """Complex CSR + CSC (values complex64 or complex128)."""
        arrs = [np.ascontiguousarray(a, t) for a, t in
                ((csr_ptr, np.int32), (csr_ind, np.int32), (csr_val, np.complex128),
                 (csc_ptr, np.int32), (csc_ind, np.int32), (csc_val, np.complex128))]
        dp = C.POINTER(C.c_double)
        h = lib().tv3o_op_sparse_c(m, n, int(arrs[1].size), _i(arrs[0]), _i(arrs[1]),
                                   arrs[2].view(np.float64).ctypes.data_as(dp),
                                   _i(arrs[3]), _i(arrs[4]),
                                   arrs[5].view(np.float64).ctypes.data_as(dp))
        return cls(h, m, n, arrs)

    @classmethod
    def geometry(cls, geo):
        s = geo.struct()
        h = lib().tv3o_op_geometry(C.byref(s))
        return cls(h, geo.m, geo.n, [geo, s])

    @property
    def nnz(self):
        return lib().tv3o_op_nnz(self.h)

    def walk(self):
        ptr = np.zeros(self.m + 1, np.int64)
        lin = np.zeros(self.nnz, np.int32)
        rc = lib().tv3o_op_walk(self.h, ptr.ctypes.data_as(_LP), _i(lin))
        if rc:
            raise ValueError("not a geometry operator")
        return ptr, lin

    def gemv_f64(self, x, trans=False):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.n if trans else self.m, np.float64)
        dp = C.POINTER(C.c_double)
        lib().tv3o_gemv_f64(self.h, int(bool(trans)), x.ctypes.data_as(dp),
                            y.ctypes.data_as(dp))
        return y

    def gemv_c(self, x, trans=False):
        """y = A x or A^H x on complex64 / complex128 vectors (by x's type)."""
        x = np.asarray(x)
        t = np.complex128 if x.dtype == np.complex128 else np.complex64
        x = np.ascontiguousarray(x, t)
        y = np.zeros(self.n if trans else self.m, t)
        c128 = t == np.complex128
        fn = lib().tv3o_gemv_c128 if c128 else lib().tv3o_gemv_c64
        rt = np.float64 if c128 else np.float32
        P_ = C.POINTER(C.c_double if c128 else C.c_float)
        fn(self.h, int(bool(trans)), x.view(rt).ctypes.data_as(P_), y.view(rt).ctypes.data_as(P_))
        return y

    def AtA_blocked(self, U, submat_mb=5):
        """(Au, g) of get_g_mult's blocked dense product."""
        U = _c32(U)
        Au = np.zeros(self.m, np.float32)
        g = np.zeros(self.n, np.float32)
        if lib().tv3o_AtA_blocked(self.h, _f(U), int(submat_mb), _f(Au), _f(g)):
            raise ValueError("not a dense operator")
        return Au, g

    def gemv(self, x, trans=False):
        x = _c32(x)
        y = np.zeros(self.n if trans else self.m, np.float32)
        lib().tv3o_gemv(self.h, int(bool(trans)), _f(x), _f(y))
        return y


def D(U, dims):
    q, p, r = dims
    U = _c32(U)
    n = U.size
    Ux, Uy = np.zeros(n, np.float32), np.zeros(n, np.float32)
    Uz = np.zeros(n, np.float32) if r > 1 else None
    lib().tv3o_D(q, p, r, _f(U), _f(Ux), _f(Uy), _f(Uz))
    return Ux, Uy, Uz


def Dt(X, Y, Z, dims, exact=False):
    q, p, r = dims
    X, Y = _c32(X), _c32(Y)
    Z = _c32(Z) if r > 1 else None
    res = np.zeros(X.size, np.float32)
    lib().tv3o_Dt(q, p, r, int(bool(exact)), _f(X), _f(Y), _f(Z), _f(res))
    return res


def shrink(Ux, Uy, Uz, sx, sy, sz, dims, beta, iso=False, scal_double=False):
    q, p, r = dims
    n = q * p * r
    ins = [_c32(a) if a is not None else None for a in (Ux, Uy, Uz, sx, sy, sz)]
    Wx, Wy = np.zeros(n, np.float32), np.zeros(n, np.float32)
    Wz = np.zeros(n, np.float32) if r > 1 else None
    lam1 = lib().tv3o_shrink(q, p, r, float(beta), int(bool(iso)),
                             int(bool(scal_double)), *[_f(a) for a in ins],
                             _f(Wx), _f(Wy), _f(Wz))
    return Wx, Wy, Wz, lam1


def dot(x, y):
    x, y = _c32(x), _c32(y)
    return lib().tv3o_dot(x.size, _f(x), _f(y))


def _check_layers(op, n, blen):
    """U holds n / op.n layers (2-D dense or sparse multi-layer mode), and b
    as many measurement vectors."""
    if op.n <= 0 or n % op.n or blen != op.m * (n // op.n):
        raise ValueError(f"oracle solve: U has {n} pixels and b {blen} entries "
                         f"for an operator of {op.m} x {op.n}")


def solve(op, b, dims, opts, Ut=None, trace_cap=0):
    """Run the restated reference solver.  Returns (U, info, trace list)."""
    from tval3_amd.abi import Info, TraceRec
    q, p, r = dims
    b = _c32(b)
    _check_layers(op, q * p * r, b.size)
    U = np.zeros(q * p * r, np.float32)
    info = Info()
    trace = (TraceRec * trace_cap)() if trace_cap else None
    Utc = _c32(Ut) if Ut is not None else None
    rc = lib().tv3o_solve(op.h, C.byref(opts), q, p, r, _f(b), _f(U), _f(Utc),
                          C.byref(info), trace, int(trace_cap))
    if rc:
        raise ValueError("oracle solve: dimension mismatch")
    n_tr = min(info.total_iters, trace_cap)
    return U, info, ([trace[i] for i in range(n_tr)] if trace else [])


def solve_c(op, b, dims, opts, Ut=None, trace_cap=0):
    """The complex instantiations of the restated solver: complex64 b gives
    the complex<float>/float solve, complex128 the complex<double>/double
    one.  Returns (U, info, trace list)."""
    from tval3_amd.abi import Info, TraceRec
    q, p, r = dims
    b = np.asarray(b)
    t = np.complex128 if b.dtype == np.complex128 else np.complex64
    rt = np.float64 if t == np.complex128 else np.float32
    P_ = C.POINTER(C.c_double if t == np.complex128 else C.c_float)
    b = np.ascontiguousarray(b, t)
    _check_layers(op, q * p * r, b.size)
    U = np.zeros(q * p * r, t)
    info = Info()
    trace = (TraceRec * trace_cap)() if trace_cap else None
    Utc = np.ascontiguousarray(Ut, t) if Ut is not None else None
    fn = lib().tv3o_solve_c128 if t == np.complex128 else lib().tv3o_solve_c64
    rc = fn(op.h, C.byref(opts), q, p, r, b.view(rt).ctypes.data_as(P_),
            U.view(rt).ctypes.data_as(P_),
            Utc.view(rt).ctypes.data_as(P_) if Utc is not None else None,
            C.byref(info), trace, int(trace_cap))
    if rc:
        raise ValueError({1: "oracle solve: dimension mismatch",
                          2: "oracle solve: isotropic TV takes real data",
                          3: "oracle solve: the operator has no complex entries"}[rc])
    n_tr = min(info.total_iters, trace_cap)
    return U, info, ([trace[i] for i in range(n_tr)] if trace else [])


def shrink_c(Ux, Uy, Uz, sx, sy, sz, dims, beta, wz_fix=False):
    """Complex shrinkage (complex64): (Wx, Wy, Wz, lam1)."""
    q, p, r = dims
    c = [np.ascontiguousarray(a, np.complex64) if a is not None else None
         for a in (Ux, Uy, Uz, sx, sy, sz)]
    outs = [np.zeros_like(c[0]), np.zeros_like(c[0]),
            np.zeros_like(c[0]) if r > 1 else None]
    fp = lambda a: a.view(np.float32).ctypes.data_as(_FP) if a is not None else None  # noqa: E731
    lam1 = lib().tv3o_shrink_c64(q, p, r, float(beta), int(bool(wz_fix)),
                                 *[fp(a) for a in c + outs])
    return outs[0], outs[1], outs[2], lam1


def solve_f64(op, b, dims, opts, Ut=None, trace_cap=0):
    """The double/double instantiation of the restated solver (FP64 vectors
    and scalars): the FP64 reference for reduced-precision results."""
    from tval3_amd.abi import Info, TraceRec
    q, p, r = dims
    dp = C.POINTER(C.c_double)
    b = np.ascontiguousarray(b, np.float64)
    _check_layers(op, q * p * r, b.size)
    U = np.zeros(q * p * r, np.float64)
    info = Info()
    trace = (TraceRec * trace_cap)() if trace_cap else None
    Utc = np.ascontiguousarray(Ut, np.float64) if Ut is not None else None
    rc = lib().tv3o_solve_f64(op.h, C.byref(opts), q, p, r, b.ctypes.data_as(dp),
                              U.ctypes.data_as(dp),
                              Utc.ctypes.data_as(dp) if Utc is not None else None,
                              C.byref(info), trace, int(trace_cap))
    if rc:
        raise ValueError("oracle solve: dimension mismatch")
    n_tr = min(info.total_iters, trace_cap)
    return U, info, ([trace[i] for i in range(n_tr)] if trace else [])
