"""Operator files: geometry and sparse-matrix descriptions on disk, through the
library's tv3_geometry_save/load and tv3_sparse_save/load (host only, no GPU).

The reference fills its geometry and sparse_mat<T> containers in the caller
(3d/mat_mult.h:148-202, 2d/mat_mult.h:38-91); these files hold the same
fields so that a description is written once and planned later.
"""
import ctypes as C

import numpy as np

from . import lib, Tval3Error
from .abi import Geometry
from .synth import GeometryArrays

_VP = C.c_void_p
_I = C.c_int


def _check(rc, what):
    if rc:
        raise Tval3Error(rc, lib().tv3_io_error().decode())


def save_geometry(path, geo):
    """Write a GeometryArrays to `path`."""
    g = geo.struct()
    _check(lib().tv3_geometry_save(None, str(path).encode(), C.byref(g)), "tv3_geometry_save")


def load_geometry(path):
    """Read a geometry file into a GeometryArrays (copies of the arrays)."""
    g = Geometry()
    st = _VP()
    _check(lib().tv3_geometry_load(None, str(path).encode(), C.byref(g), C.byref(st)),
           f"tv3_geometry_load({path})")
    try:
        P, E, R = g.num_mps, g.num_emitters, g.num_receivers

        def arr(ptr, n):
            return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

        return GeometryArrays(P, arr(g.use_path, E * R),
                              arr(g.x_emitters, P * E), arr(g.y_emitters, P * E),
                              arr(g.z_emitters, P * E), arr(g.x_receivers, P * R),
                              arr(g.y_receivers, P * R), arr(g.z_receivers, P * R),
                              (g.rv_x, g.rv_y, g.rv_z), g.num_paths, g.scale_factor)
    finally:
        lib().tv3_storage_free(st)


def save_sparse(path, m, n, csr_ptr, csr_ind, csr_val, csc_ptr, csc_ind, csc_val):
    """Write CSR and CSC arrays of an m x n matrix to `path`."""
    ints = [np.ascontiguousarray(a, np.int32) for a in (csr_ptr, csr_ind, csc_ptr, csc_ind)]
    vals = [np.ascontiguousarray(a, np.float32) for a in (csr_val, csc_val)]
    rp, ri, cp, ci = ints
    rv, cv = vals
    if rp.size != m + 1 or cp.size != n + 1 or not (ri.size == rv.size == ci.size == cv.size):
        raise Tval3Error(1, "Argument error: sparse array sizes do not match m, n, nnz")
    p = lambda a: a.ctypes.data_as(_VP)  # noqa: E731
    _check(lib().tv3_sparse_save(None, str(path).encode(), int(m), int(n), int(ri.size),
                                 p(rp), p(ri), p(rv), p(cp), p(ci), p(cv)), "tv3_sparse_save")


def load_sparse(path):
    """Read a sparse-matrix file: (m, n, csr_ptr, csr_ind, csr_val, csc_ptr,
    csc_ind, csc_val) as numpy copies."""
    m, n, z = _I(), _I(), _I()
    ptrs = [_VP() for _ in range(6)]
    st = _VP()
    _check(lib().tv3_sparse_load(None, str(path).encode(), C.byref(m), C.byref(n), C.byref(z),
                                 *[C.byref(q) for q in ptrs], C.byref(st)),
           f"tv3_sparse_load({path})")
    try:
        M, N, Z = m.value, n.value, z.value

        def arr(q, count, t):
            if count == 0:
                return np.zeros(0, t)
            ct = C.c_int if t == np.int32 else C.c_float
            return np.ctypeslib.as_array(C.cast(q, C.POINTER(ct)), shape=(count,)).copy()

        return (M, N, arr(ptrs[0], M + 1, np.int32), arr(ptrs[1], Z, np.int32),
                arr(ptrs[2], Z, np.float32), arr(ptrs[3], N + 1, np.int32),
                arr(ptrs[4], Z, np.int32), arr(ptrs[5], Z, np.float32))
    finally:
        lib().tv3_storage_free(st)
