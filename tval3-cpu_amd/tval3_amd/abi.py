"""ctypes mirrors of the POD structs declared in include/tval3_amd.h.

The field order and types follow the C header exactly; any change there must be
made here too (tests/test_abi.py checks the sizes against the compiled library).
"""
import ctypes as C

__all__ = ["Options", "Info", "TraceRec", "Geometry", "ProfileSlot", "default_options"]


class Options(C.Structure):
    """tv3_options: the reference's tval3_options<T> fields plus engine settings."""
    _fields_ = [
        ("mu", C.c_double), ("beta", C.c_double),
        ("mu0", C.c_double), ("beta0", C.c_double),
        ("tol", C.c_double), ("tol_inn", C.c_double),
        ("maxit", C.c_int), ("maxcnt", C.c_int),
        ("gam", C.c_double), ("gamma", C.c_double),
        ("rate_gam", C.c_double), ("rate_cnt", C.c_double),
        ("c", C.c_double),
        ("nonneg", C.c_int), ("isreal", C.c_int),
        ("numInitIterations", C.c_uint),
        ("skipMulRatio", C.c_double), ("maxRelativeChange", C.c_double),
        ("scal_double", C.c_int), ("tv_iso", C.c_int), ("dt_exact", C.c_int),
        ("ls_exact", C.c_int), ("graph", C.c_int), ("submat_mb", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Info(C.Structure):
    """tv3_info: the reference's tval3_info<T>."""
    _fields_ = [("secs", C.c_double), ("outer_iters", C.c_int),
                ("total_iters", C.c_int), ("rel_chg", C.c_double),
                ("rel_error", C.c_double)]


class TraceRec(C.Structure):
    """tv3_trace_rec: one record per inner iteration."""
    _fields_ = [("iter", C.c_int), ("outer", C.c_int),
                ("backtracks", C.c_int), ("fallback", C.c_int),
                ("tau", C.c_double), ("alpha", C.c_double), ("f", C.c_double),
                ("C", C.c_double), ("rel_chg", C.c_double),
                ("lam1", C.c_double), ("lam2", C.c_double),
                ("lam3", C.c_double), ("lam4", C.c_double),
                ("lam5", C.c_double), ("mu", C.c_double),
                ("beta", C.c_double)]


class ProfileSlot(C.Structure):
    """tv3_profile_slot: one slot of the launch-timing report."""
    _fields_ = [("slot", C.c_int), ("runs", C.c_longlong), ("ms", C.c_double),
                ("name", C.c_char * 96)]


_IP = C.POINTER(C.c_int)


class Geometry(C.Structure):
    """tv3_geometry: the reference's `geometry` (live 3-D projector form)."""
    _fields_ = [("num_mps", C.c_int), ("num_emitters", C.c_int),
                ("num_receivers", C.c_int), ("num_paths", C.c_int),
                ("use_path", _IP),
                ("x_emitters", _IP), ("y_emitters", _IP), ("z_emitters", _IP),
                ("x_receivers", _IP), ("y_receivers", _IP),
                ("z_receivers", _IP),
                ("rv_x", C.c_int), ("rv_y", C.c_int), ("rv_z", C.c_int),
                ("scale_factor", C.c_float)]


def default_options(**kw):
    """The public TVAL3 defaults (the reference's own defaults are in a header it
    does not ship), overridden by keyword arguments."""
    o = Options()
    vals = dict(mu=2.0 ** 8, beta=2.0 ** 5, mu0=2.0 ** 8, beta0=2.0 ** 5,
                tol=1e-4, tol_inn=1e-3, maxit=300, maxcnt=10, gam=0.9995,
                gamma=0.6, rate_gam=0.9, rate_cnt=2.0, c=1e-5, nonneg=0,
                isreal=1, numInitIterations=0, skipMulRatio=0.0,
                maxRelativeChange=0.0, scal_double=0, tv_iso=0, dt_exact=0,
                ls_exact=0, graph=0, submat_mb=5)
    vals.update(kw)
    for k, v in vals.items():
        setattr(o, k, v)
    return o
