"""tval3_amd — Python binding of the MI355X TVAL3 engine (libtval3_amd.so).

A thin ctypes layer over the C-ABI declared in include/tval3_amd.h.  PyTorch is
used only to own device memory and to supply the HIP stream; every computation
runs in the hand-written HIP kernels of the native library.  There is no CPU
fallback: if the library is missing or no GPU is visible, the calls raise.

Reference interface mirrored: tval3_cpu_2d / tval3_cpu_3d
(2d/tval3_cpu_2d.cpp:889-1016, 3d/tval3_cpu_3d.cpp:1384-1580) and the gemv
overloads of 2d/mat_mult.h / 3d/mat_mult.h.  Argument errors raise
Tval3Error with the library's message, as the reference throws
tval3_exception.
"""
import ctypes as C
import os

import numpy as np

from .abi import Options, Info, TraceRec, Geometry, ProfileSlot, default_options  # noqa: F401

__all__ = ["Tval3Error", "lib", "lib_path", "Context", "Plan", "Options",
           "Info", "TraceRec", "default_options"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libtval3_amd.so"))
_lib = None

_VP = C.c_void_p
_I = C.c_int
_LL = C.c_longlong


class Tval3Error(RuntimeError):
    """Raised for a non-zero status of the native library (the reference's
    tval3_exception)."""

    def __init__(self, code, msg):
        super().__init__(f"[tval3 status {code}] {msg}")
        self.code = code


def lib_path():
    return _LIB_PATH


def lib():
    """Load libtval3_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"native library not built: {_LIB_PATH} "
                           "(run __graft_entry__.build())")
    L = C.CDLL(_LIB_PATH)
    sig = {
        "tv3_options_default": (None, [_VP]),
        "tv3_ctx_create": (_I, [_I, _VP, C.POINTER(_VP)]),
        "tv3_ctx_destroy": (None, [_VP]),
        "tv3_ctx_error": (C.c_char_p, [_VP]),
        "tv3_ctx_sync": (_I, [_VP]),
        "tv3_ctx_stream": (_VP, [_VP]),
        "tv3_ctx_profile": (_I, [_VP, _I]),
        "tv3_ctx_profile_read": (_I, [_VP, C.POINTER(C.c_double), C.POINTER(_LL),
                                      C.POINTER(C.c_double)]),
        "tv3_ctx_profile_slots": (_I, [_VP, C.POINTER(ProfileSlot), _I, C.POINTER(_I)]),
        "tv3_ctx_profile_lanes": (_I, [_VP, C.POINTER(_LL)]),
        "tv3_plan_dense": (_I, [_VP, _I, _I, _I, _VP, C.POINTER(_VP)]),
        "tv3_plan_sparse": (_I, [_VP, _I, _I, _I] + [_VP] * 6 + [C.POINTER(_VP)]),
        "tv3_plan_geometry": (_I, [_VP, _VP, C.POINTER(_VP)]),
        "tv3_plan_destroy": (None, [_VP]),
        "tv3_plan_shape": (_I, [_VP, C.POINTER(_LL), C.POINTER(_LL), C.POINTER(_LL)]),
        "tv3_plan_walk": (_I, [_VP, _VP, _VP, _VP]),
        "tv3_gemv": (_I, [_VP, _VP, _I, _VP, _VP]),
        "tv3_gemv_noalloc": (_I, [_VP, _VP, _I, _VP, _VP]),
        "tv3_D": (_I, [_VP, _I, _I, _I, _VP, _VP, _VP, _VP]),
        "tv3_Dt": (_I, [_VP, _I, _I, _I, _I, _VP, _VP, _VP, _VP]),
        "tv3_shrink": (_I, [_VP, _I, _I, _I, C.c_double, _I, _I] + [_VP] * 9
                       + [C.POINTER(C.c_double)]),
        "tv3_dot": (_I, [_VP, _LL, _VP, _VP, C.POINTER(C.c_double)]),
        "tv3_nrm2": (_I, [_VP, _LL, _VP, C.POINTER(C.c_double)]),
        "tv3_solve": (_I, [_VP, _VP, _VP, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _I]),
        "tv3_solve_batch": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
        "tv3_plan_sparse_f64": (_I, [_VP, _I, _I, _I] + [_VP] * 6 + [C.POINTER(_VP)]),
        "tv3_gemv_f64": (_I, [_VP, _VP, _I, _VP, _VP]),
        "tv3_solve_f64": (_I, [_VP, _VP, _VP, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _I]),
        "tv3_solve_batch_f64": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
        "tv3_plan_sparse_c64": (_I, [_VP, _I, _I, _I] + [_VP] * 6 + [C.POINTER(_VP)]),
        "tv3_plan_sparse_c128": (_I, [_VP, _I, _I, _I] + [_VP] * 6 + [C.POINTER(_VP)]),
        "tv3_gemv_c64": (_I, [_VP, _VP, _I, _VP, _VP]),
        "tv3_gemv_c128": (_I, [_VP, _VP, _I, _VP, _VP]),
        "tv3_solve_c64": (_I, [_VP, _VP, _VP, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _I]),
        "tv3_solve_c128": (_I, [_VP, _VP, _VP, _I, _I, _I, _VP, _VP, _VP, _VP, _VP, _I]),
        "tv3_solve_batch_c64": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
        "tv3_solve_batch_c128": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
        "tv3_mem_alloc": (_I, [_VP, C.c_size_t, C.POINTER(_VP)]),
        "tv3_mem_free": (_I, [_VP, _VP]),
        "tv3_mem_upload": (_I, [_VP, _VP, _VP, C.c_size_t]),
        "tv3_mem_download": (_I, [_VP, _VP, _VP, C.c_size_t]),
        "tv3_unrolled_forward": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _VP, _VP, _VP, _VP,
                                      C.POINTER(_VP)]),
        "tv3_unrolled_backward": (_I, [_VP, _VP, _VP, _VP, C.POINTER(C.c_double)]),
        "tv3_tape_destroy": (None, [_VP]),
        "tv3_geometry_save": (_I, [_VP, C.c_char_p, _VP]),
        "tv3_geometry_load": (_I, [_VP, C.c_char_p, _VP, C.POINTER(_VP)]),
        "tv3_sparse_save": (_I, [_VP, C.c_char_p, _I, _I, _I] + [_VP] * 6),
        "tv3_sparse_load": (_I, [_VP, C.c_char_p] + [C.POINTER(_I)] * 3
                            + [C.POINTER(_VP)] * 7),
        "tv3_storage_free": (None, [_VP]),
        "tv3_io_error": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(t):
    return None if t is None else _VP(t.data_ptr())


def _vkind(x, f64):
    """Vector kind of an entry point: 'f32' / 'f64' for real data, 'c64' /
    'c128' for complex data (complex128 data or f64=True: 'c128')."""
    cplx = x is not None and (np.iscomplexobj(x) if not hasattr(x, "is_complex") else x.is_complex())
    wide = f64 or (x is not None and str(getattr(x, "dtype", "")).endswith("complex128"))
    if cplx:
        return "c128" if wide else "c64"
    return "f64" if f64 else "f32"


_SUFFIX = {"f32": "", "f64": "_f64", "c64": "_c64", "c128": "_c128"}


class Context:
    """A device context: HIP stream + device workspace (tv3_ctx)."""

    def __init__(self, device=0, stream=None):
        import torch
        self.torch = torch
        if not torch.cuda.is_available():
            raise RuntimeError("tval3_amd needs a HIP device")
        self.device = torch.device("cuda", device)
        if stream is None:
            # a stream of our own (default priority); every native call is
            # ordered after torch's current stream and torch's current stream
            # after it (_enter/_exit), so tensors move freely between the two
            stream = torch.cuda.Stream(self.device)
        self.stream = stream
        h = _VP()
        rc = lib().tv3_ctx_create(device, _VP(stream.cuda_stream), C.byref(h))
        if rc:
            raise Tval3Error(rc, "tv3_ctx_create failed")
        self.h = h
        import weakref
        self._plans = weakref.WeakSet()

    def close(self):
        if getattr(self, "h", None):
            for p in list(self._plans):
                p.close()
            lib().tv3_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise Tval3Error(rc, lib().tv3_ctx_error(self.h).decode())

    def _enter(self):
        self.stream.wait_stream(self.torch.cuda.current_stream(self.device))

    def _exit(self):
        self.torch.cuda.current_stream(self.device).wait_stream(self.stream)

    def _call(self, fn, *args):
        """Runs one native entry point ordered with torch's current stream."""
        self._enter()
        rc = fn(*args)
        self._exit()
        self._check(rc)

    def sync(self):
        self._check(lib().tv3_ctx_sync(self.h))

    def profile(self, enable=True):
        """Start (and reset) or stop per-launch timing: False / 0 off, True / 1
        the A / A^T products only, 2 every launch of the solver (slower), 3 the
        products only without the device-side byte and lane counts of batched
        launches."""
        level = int(enable)
        if level not in (0, 1, 2, 3):
            raise ValueError(f"profile level {enable!r}: 0, 1, 2 or 3")
        self._check(lib().tv3_ctx_profile(self.h, level))

    def profile_slots(self):
        """[(slot, name, runs, ms_total)] of the slots that ran since
        profile() was switched on, in slot order (slot numbers as in the
        reference's profile_info table where one pass implements that
        function, 30 and up for fused passes)."""
        n = C.c_int(0)
        self._check(lib().tv3_ctx_profile_slots(self.h, None, 0, C.byref(n)))
        buf = (ProfileSlot * max(1, n.value))()
        self._check(lib().tv3_ctx_profile_slots(self.h, buf, n.value, C.byref(n)))
        return [(r.slot, r.name.decode(), r.runs, r.ms) for r in buf[:n.value]]

    def profile_report(self):
        """The slot table as text, one line per slot in the reference's
        PROFILING print format (name, ms per run, runs, total ms;
        3d/tval3_cpu_3d.cpp:1337-1348)."""
        return "".join(f"{name}:\t {ms / runs:.2f} ms (runs: {runs}, total: {ms:.1f} ms)\n"
                       for _, name, runs, ms in self.profile_slots() if runs)

    def profile_read(self):
        """{'A': (ms_total, launches, bytes, lanes), 'At': (...)}: summed
        launch milliseconds, launch count, summed compulsory bytes (operator
        once per computing group, vectors per computing instance) and the
        number of instance products the launches computed."""
        ms = (C.c_double * 2)()
        cnt = (_LL * 2)()
        by = (C.c_double * 2)()
        self._check(lib().tv3_ctx_profile_read(self.h, ms, cnt, by))
        ln = (_LL * 2)()
        self._check(lib().tv3_ctx_profile_lanes(self.h, ln))
        return {"A": (ms[0], cnt[0], by[0], ln[0]), "At": (ms[1], cnt[1], by[1], ln[1])}

    def tensor(self, a, dtype=None):
        """Host array -> contiguous device tensor (float32 unless given)."""
        t = self.torch
        if isinstance(a, t.Tensor):
            x = a.to(self.device)
        else:
            x = t.from_numpy(np.ascontiguousarray(a))
            x = x.to(self.device)
        if dtype is not None:
            x = x.to(dtype)
        return x.contiguous()

    # ---- plans -----------------------------------------------------------
    def plan_dense(self, A, bf16=False, f64=False):
        """Dense A on the device: float32 (default), bfloat16 (bf16=True) or,
        for the double-vector entry points, float64 (f64=True, mat<double>).
        Complex A gives a complex plan (complex64; complex128 when A is
        complex128 or f64=True) for the complex entry points."""
        t = self.torch
        if bf16 and f64:
            raise Tval3Error(1, "Argument error: bf16 and f64 exclude each other")
        k = _vkind(A, f64)
        if k in ("c64", "c128") and bf16:
            raise Tval3Error(1, "Argument error: bf16 takes real A")
        dt = {"c64": (t.complex64, 3), "c128": (t.complex128, 4), "f64": (t.float64, 2),
              "f32": (t.bfloat16, 1) if bf16 else (t.float32, 0)}[k]
        At = self.tensor(A, dt[0])
        m, n = At.shape
        h = _VP()
        self._call(lib().tv3_plan_dense, self.h, m, n, dt[1], _ptr(At), C.byref(h))
        return Plan(self, h, [At])

    def plan_sparse(self, m, n, csr_ptr, csr_ind, csr_val, csc_ptr, csc_ind,
                    csc_val, f64=False):
        """CSR + CSC copies of one m x n matrix; f64=True: double values
        (sparse_mat<double>, for the double-vector entry points); complex
        values: a complex plan (complex128 values or f64=True: complex128)."""
        t = self.torch
        k = _vkind(csr_val, f64)
        vt = {"f32": t.float32, "f64": t.float64, "c64": t.complex64, "c128": t.complex128}[k]
        arrs = [self.tensor(csr_ptr, t.int32), self.tensor(csr_ind, t.int32),
                self.tensor(csr_val, vt), self.tensor(csc_ptr, t.int32),
                self.tensor(csc_ind, t.int32), self.tensor(csc_val, vt)]
        h = _VP()
        fn = getattr(lib(), "tv3_plan_sparse" + _SUFFIX[k])
        self._call(fn, self.h, m, n, int(arrs[1].numel()), *[_ptr(a) for a in arrs], C.byref(h))
        return Plan(self, h, arrs)

    def plan_geometry(self, geo):
        s = geo.struct()
        h = _VP()
        self._call(lib().tv3_plan_geometry, self.h, C.byref(s), C.byref(h))
        return Plan(self, h, [])

    # ---- operators ------------------------------------------------------
    def gemv(self, plan, x, trans=False, f64=False):
        """y = A x (A^T x, complex: A^H x, with trans); f64=True: double
        vectors (tv3_gemv_f64); complex x: tv3_gemv_c64 / _c128."""
        t = self.torch
        k = _vkind(x, f64)
        vt = {"f32": t.float32, "f64": t.float64, "c64": t.complex64, "c128": t.complex128}[k]
        x = self.tensor(x, vt)
        y = t.empty(plan.n if trans else plan.m, dtype=vt, device=self.device)
        fn = getattr(lib(), "tv3_gemv" + _SUFFIX[k])
        self._call(fn, self.h, plan.h, int(bool(trans)), _ptr(x), _ptr(y))
        return y

    def D(self, U, dims):
        t = self.torch
        q, p, r = dims
        U = self.tensor(U, t.float32)
        out = [t.empty_like(U), t.empty_like(U)] + ([t.empty_like(U)] if r > 1 else [None])
        self._call(lib().tv3_D, self.h, q, p, r, _ptr(U), *[_ptr(o) for o in out])
        return out

    def Dt(self, X, Y, Z, dims, exact=False):
        t = self.torch
        q, p, r = dims
        X, Y = self.tensor(X, t.float32), self.tensor(Y, t.float32)
        Z = self.tensor(Z, t.float32) if r > 1 else None
        res = t.empty_like(X)
        self._call(lib().tv3_Dt, self.h, q, p, r, int(bool(exact)), _ptr(X),
                                 _ptr(Y), _ptr(Z), _ptr(res))
        return res

    def shrink(self, Ux, Uy, Uz, sx, sy, sz, dims, beta, iso=False,
               scal_double=False):
        t = self.torch
        q, p, r = dims
        ins = [self.tensor(a, t.float32) if a is not None else None
               for a in (Ux, Uy, Uz, sx, sy, sz)]
        outs = [t.empty_like(ins[0]), t.empty_like(ins[0]),
                t.empty_like(ins[0]) if r > 1 else None]
        lam1 = C.c_double()
        self._call(lib().tv3_shrink, self.h, q, p, r, float(beta), int(bool(iso)),
                                     int(bool(scal_double)),
                                     *[_ptr(a) for a in ins + outs],
                                     C.byref(lam1))
        return outs[0], outs[1], outs[2], lam1.value

    def dot(self, x, y):
        t = self.torch
        x, y = self.tensor(x, t.float32), self.tensor(y, t.float32)
        out = C.c_double()
        self._call(lib().tv3_dot, self.h, x.numel(), _ptr(x), _ptr(y), C.byref(out))
        return out.value

    # ---- solver -----------------------------------------------------------
    def solve(self, plan, b, dims, opts=None, Ut=None, trace_cap=0, f64=False):
        """tval3_cpu_2d/3d on the device: returns (U tensor, Info, [TraceRec]).
        f64=True: the double/double solve (tv3_solve_f64, float64 tensors);
        complex b: the complex solve (tv3_solve_c64, or _c128 for complex128
        b or f64=True)."""
        t = self.torch
        if opts is None:
            opts = default_options()
        q, p, r = dims
        k = _vkind(b, f64)
        vt = {"f32": t.float32, "f64": t.float64, "c64": t.complex64, "c128": t.complex128}[k]
        b = self.tensor(b, vt)
        if b.numel() != plan.m * _layers(plan, dims):
            raise Tval3Error(1, "Argument error: b.len must be equal to A.dim_y * layers!")
        U = t.empty(q * p * r, dtype=vt, device=self.device)
        Utt = self.tensor(Ut, vt) if Ut is not None else None
        if Utt is not None and Utt.numel() != q * p * r:
            raise Tval3Error(1, "Argument error: U and Ut must have the same size!")
        info = Info()
        trace = (TraceRec * trace_cap)() if trace_cap else None
        fn = getattr(lib(), "tv3_solve" + _SUFFIX[k])
        self._call(fn, self.h, plan.h, C.byref(opts), q, p, r,
                                    _ptr(b), _ptr(U), _ptr(Utt), C.byref(info),
                                    trace, int(trace_cap))
        n_tr = min(info.total_iters, trace_cap)
        return U, info, ([trace[i] for i in range(n_tr)] if trace else [])

    # ---- K unrolled iterations and their adjoint (differentiation harness) ----
    def unrolled_forward(self, plan, b, dims, opts, K, ckpt, alpha, first):
        """U = u_K / scl of the unrolled map (tv3_unrolled_forward) and its
        tape.  alpha / first: the K step factors and first-step flags."""
        t = self.torch
        q, p = dims[0], dims[1]
        b = self.tensor(b, t.float32)
        if b.numel() != plan.m:
            raise Tval3Error(1, "Argument error: b.len must be equal to A.dim_y!")
        al = np.ascontiguousarray(alpha, np.float64)
        fs = np.ascontiguousarray(first, np.int32)
        if al.size != K or fs.size != K:
            raise Tval3Error(1, "Argument error: alpha and first_step need K entries")
        U = t.empty(q * p, dtype=t.float32, device=self.device)
        h = _VP()
        self._call(lib().tv3_unrolled_forward, self.h, plan.h, C.byref(opts), q, p,
                   int(K), int(ckpt), al.ctypes.data_as(_VP), fs.ctypes.data_as(_VP),
                   _ptr(b), _ptr(U), C.byref(h))
        return U, Tape(self, h, plan.m)

    def unrolled_backward(self, tape, dU):
        """(dL/db tensor, dL/dmu float) for the cotangent dU."""
        t = self.torch
        dU = self.tensor(dU, t.float32)
        db = t.empty(tape.m, dtype=t.float32, device=self.device)
        dmu = C.c_double()
        self._call(lib().tv3_unrolled_backward, self.h, tape.h, _ptr(dU), _ptr(db),
                   C.byref(dmu))
        return db, dmu.value

    def solve_batch(self, plans, bs, dims, opts=None, Uts=None, f64=False):
        """B reconstructions in one batched launch sequence.  Returns
        (list of U tensors, list of Info).  f64=True: tv3_solve_batch_f64;
        complex b: tv3_solve_batch_c64 / _c128."""
        t = self.torch
        k = _vkind(bs[0] if len(bs) else None, f64)
        vt = {"f32": t.float32, "f64": t.float64, "c64": t.complex64, "c128": t.complex128}[k]
        if opts is None:
            opts = default_options()
        B = len(plans)
        if len(bs) != B or len(dims) != B:
            raise Tval3Error(1, "Argument error: plans, b and dims differ in length")
        bts = [self.tensor(b, vt) for b in bs]
        for pl, x, d in zip(plans, bts, dims):
            if x.numel() != pl.m * _layers(pl, d):
                raise Tval3Error(1, "Argument error: b.len must be equal to A.dim_y * layers!")
        if Uts is not None:
            for d, u in zip(dims, Uts):
                if u is not None and np.size(u) != d[0] * d[1] * d[2]:
                    raise Tval3Error(1, "Argument error: U and Ut must have the same size!")
        Us = [t.empty(d[0] * d[1] * d[2], dtype=vt, device=self.device)
              for d in dims]
        Utts = ([self.tensor(u, vt) if u is not None else None for u in Uts]
                if Uts is not None else None)
        parr = (_VP * B)(*[p.h for p in plans])
        dims_arr = (C.c_int * (3 * B))(*[int(v) for d in dims for v in d])
        barr = (_VP * B)(*[x.data_ptr() for x in bts])
        uarr = (_VP * B)(*[x.data_ptr() for x in Us])
        utarr = ((_VP * B)(*[(x.data_ptr() if x is not None else None) for x in Utts])
                 if Utts is not None else None)
        infos = (Info * B)()
        fn = getattr(lib(), "tv3_solve_batch" + _SUFFIX[k])
        self._call(fn, self.h, B, parr, C.byref(opts), dims_arr, barr, uarr, utarr, infos)
        return Us, [infos[i] for i in range(B)]


def _layers(plan, dims):
    """Layers of a 2-D multi-layer solve: U's pixel count over A's columns
    (1 when they are equal; the library rejects other mismatches)."""
    n = int(dims[0]) * int(dims[1]) * int(dims[2])
    return n // plan.n if plan.n > 0 and n % plan.n == 0 and n > plan.n else 1


class Plan:
    """A measurement-operator plan (keeps the borrowed device arrays alive)."""

    def __init__(self, ctx, handle, keep):
        self.ctx = ctx
        self.h = handle
        self._keep = keep
        ctx._plans.add(self)
        m, n, nnz = _LL(), _LL(), _LL()
        lib().tv3_plan_shape(handle, C.byref(m), C.byref(n), C.byref(nnz))
        self.m, self.n, self.nnz = m.value, n.value, nnz.value

    def walk(self):
        """Voxel index list of every row as computed by the device walk kernel."""
        ptr = np.zeros(self.m + 1, np.int64)
        lin = np.zeros(max(self.nnz, 1), np.int32)
        self.ctx._call(lib().tv3_plan_walk, self.ctx.h, self.h,
                                            ptr.ctypes.data_as(_VP),
                                            lin.ctypes.data_as(_VP))
        return ptr, lin[:self.nnz]

    def close(self):
        if getattr(self, "h", None):
            lib().tv3_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tape:
    """Checkpoints of one unrolled forward pass (tv3_tape)."""

    def __init__(self, ctx, h, m):
        self.ctx, self.h, self.m = ctx, h, m

    def close(self):
        if getattr(self, "h", None):
            lib().tv3_tape_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
