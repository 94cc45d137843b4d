"""Python mirror of T-FlowS's Solvers_Mod on the GPU (libtfs_solver.so).

Same names, argument meaning and error behaviour as the reference subroutines
(Sources/Process/Solvers_Mod/{Cg,Bicg,Cgs}.f90):

    niter, ini_res, fin_res = Cg(A, x, r1, prec, niter, tol, ini_res=0.0)

* ``A``      -- ``Matrix`` (the Matrix_Type counterpart, structure resident on
                the device, values uploaded with ``set_values``);
* ``x``      -- initial guess in, solution out (in place);
* ``r1``     -- right-hand side in, final recurrence residual out (in place);
* ``prec``   -- 0 none, 1 diagonal, 2 the reference's incomplete Cholesky;
* ``niter``  -- iteration cap in; the returned count is k on convergence at
                iteration k, 0 on an early exit, cap+1 without convergence;
* ``tol``    -- ABSOLUTE tolerance on ||r||_2 (Cg.f90:174).

``x`` and ``r1`` may be numpy float64 arrays (host path, ``tfs_solve``) or
torch float64 tensors on the GPU (device path, ``tfs_solve_device``).  As in
the reference there is no exception for non-convergence; the last status is
kept in ``A.status``.  Invalid arguments and device failures raise.
There is no CPU fallback: without the HIP library this module raises.
"""
import atexit
import ctypes
import weakref

import numpy as np

from . import _native

TFS_OK = 0
TFS_NOT_CONVERGED = 1
TFS_CANCELLED = 2
TFS_ERR_INVALID = -1
TFS_ERR_NOMEM = -2
TFS_ERR_BREAKDOWN = -3
TFS_ERR_DEVICE = -4
TFS_ERR_NO_DEVICE = -5

CG, BICG, CGS = 0, 1, 2
TFS_INNER_REFERENCE, TFS_INNER_COLOURED, TFS_INNER_BRICKS, TFS_INNER_AUTO = 0, 1, 2, 3
TFS_REFINE_FCG, TFS_REFINE_DEFECT, TFS_REFINE_RELIABLE = 0, 1, 2
TFS_SWEEP_LEVELS, TFS_SWEEP_STATIC, TFS_SWEEP_DYNAMIC, TFS_SWEEP_COOPERATIVE = 0, 1, 2, 3
_METHODS = {"cg": CG, "bicg": BICG, "cgs": CGS}

_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


class SolverError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} ({status})")


class _Info(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nnz", ctypes.c_longlong),
                ("levels_lower", ctypes.c_int), ("levels_upper", ctypes.c_int),
                ("row_blocks", ctypes.c_int), ("device_bytes", ctypes.c_longlong),
                ("inner_colours", ctypes.c_int), ("spmv_rows", ctypes.c_int),
                ("spmv_waves", ctypes.c_int), ("inner_block", ctypes.c_int),
                ("inner_levels", ctypes.c_int), ("ell64_slices", ctypes.c_int),
                ("inner_order", ctypes.c_int), ("sweep_groups", ctypes.c_int),
                ("inner_sweep_groups", ctypes.c_int), ("inner_brick_edge", ctypes.c_int)]


_LIVE = []  # weak references to the library objects, in creation order


def _track(obj):
    _LIVE.append(weakref.ref(obj))


def _close_all():
    """Destroys the library objects still alive, newest first (an inner or
    dependent object before what it was made from).  Registered with atexit
    after torch is imported, so it runs before torch's exit handlers, before
    the interpreter's teardown (whose garbage collection would call the
    destroys in no particular order) and before the C runtime's exit handlers
    shut the HIP runtime down: every stream, event and allocation goes back
    while everything it belongs to is still up."""
    while _LIVE:
        obj = _LIVE.pop()()
        if obj is not None:
            try:
                obj.close()
            except Exception:  # pragma: no cover -- best effort at exit
                pass


def _lib():
    # torch first: its HIP runtime then serves this library too (same soname),
    # so device pointers from torch tensors are valid here.
    import torch  # noqa: F401
    lib = _native.solver_lib()
    if getattr(lib, "_tfs_typed", False):
        return lib
    atexit.register(_close_all)
    lib.tfs_status_string.restype = ctypes.c_char_p
    lib.tfs_status_string.argtypes = [ctypes.c_int]
    lib.tfs_init.argtypes = [ctypes.c_int]
    lib.tfs_matrix_create.argtypes = [ctypes.c_int, ctypes.c_longlong, _I32P, _I32P, _I32P,
                                      ctypes.c_int, ctypes.POINTER(_VP)]
    lib.tfs_matrix_destroy.argtypes = [_VP]
    lib.tfs_matrix_set_values.argtypes = [_VP, _F64P]
    lib.tfs_matrix_set_values_device.argtypes = [_VP, _VP]
    lib.tfs_matrix_values_changed.argtypes = [_VP]
    lib.tfs_matrix_values_device.restype = _VP
    lib.tfs_matrix_values_device.argtypes = [_VP]
    lib.tfs_matrix_get_info.argtypes = [_VP, ctypes.POINTER(_Info)]
    lib.tfs_solve.argtypes = [_VP, ctypes.c_int, ctypes.c_int, _F64P, _F64P, _IP,
                              ctypes.c_double, _DP, _DP]
    lib.tfs_solve_device.argtypes = [_VP, ctypes.c_int, ctypes.c_int, _VP, _VP, _IP,
                                     ctypes.c_double, _DP, _DP]
    lib.tfs_spmv_device.argtypes = [_VP, _VP, _VP]
    lib.tfs_residual_device.argtypes = [_VP, _VP, _VP]
    lib.tfs_prec_form.argtypes = [_VP, ctypes.c_int]
    lib.tfs_prec_solve_device.argtypes = [_VP, ctypes.c_int, _VP, _VP]
    lib.tfs_prec_diag_device.argtypes = [_VP, _VP]
    lib.tfs_spmv_time.argtypes = [_VP, _VP, _VP, ctypes.c_int, _DP]
    lib.tfs_sweep_time.argtypes = [_VP, ctypes.c_int, _DP]
    lib.tfs_spmv_time_cold.argtypes = [_VP, _VP, _VP, ctypes.c_int, _DP]
    lib.tfs_solve_mixed_device.argtypes = [_VP, ctypes.c_int, _VP, _VP, _IP, _IP,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           _DP, _DP]
    lib.tfs_prec_diag32_device.argtypes = [_VP, ctypes.c_int, _VP]
    lib.tfs_matrix_set_inner_order.argtypes = [_VP, ctypes.c_int]
    lib.tfs_matrix_get_values.argtypes = [_VP, _F64P]
    lib.tfs_matrix_set_refinement.argtypes = [_VP, ctypes.c_int]
    lib.tfs_matrix_set_inner_block.argtypes = [_VP, ctypes.c_int]
    lib.tfs_matrix_tune_spmv.argtypes = [_VP, _IP, _IP]
    lib.tfs_matrix_set_spmv_shape.argtypes = [_VP, ctypes.c_int, ctypes.c_int]
    lib.tfs_matrix_set_sweep_schedule.argtypes = [_VP, ctypes.c_int]
    lib.tfs_timing_start.argtypes = [ctypes.c_int]
    lib.tfs_timing_stop.argtypes = [_DP, ctypes.POINTER(ctypes.c_longlong)]
    lib._tfs_typed = True
    return lib


def status_string(status):
    return _lib().tfs_status_string(status).decode()


def _check(rc, what):
    if rc < 0 and rc != TFS_ERR_BREAKDOWN:
        raise SolverError(rc, what)
    return rc


def _is_device_tensor(a):
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(a, torch.Tensor) and a.is_cuda


def _dptr(t, n, what):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
            and t.is_contiguous() and t.numel() >= n):
        raise ValueError(f"{what}: expected a contiguous float64 GPU tensor of >= {n} entries")
    return t.data_ptr()


def _check_count(t, n, what, exact=True):
    """Size check before any device access: a per-face or per-cell array
    whose length differs from the mesh's would make the kernels read or write
    past its end (or with the wrong stride)."""
    if t is None:
        return
    m = t.numel() if hasattr(t, "numel") else np.asarray(t).size
    if (m != n) if exact else (m < n):
        raise ValueError(f"{what}: {m} entries, the mesh needs {'exactly ' if exact else ''}{n}")


def _sync():
    import torch
    torch.cuda.synchronize()


class Matrix:
    """Matrix_Type counterpart (Sources/Process/Matrix_Mod.f90:36-49) on the GPU."""

    def __init__(self, row, col, dia, val=None, base=0):
        lib = _lib()
        row = np.ascontiguousarray(row, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        dia = np.ascontiguousarray(dia, dtype=np.int32)
        n = len(row) - 1
        nnz = int(row[-1]) - base if n >= 0 else 0
        h = _VP()
        _check(lib.tfs_matrix_create(n, nnz, row, col, dia, base, ctypes.byref(h)),
               "tfs_matrix_create")
        self._h = h
        _track(self)
        self._lib = lib
        self.n = n
        self.nnz = nnz
        self.status = TFS_OK
        if val is not None:
            self.set_values(val)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tfs_matrix_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self):
        i = _Info()
        _check(self._lib.tfs_matrix_get_info(self._h, ctypes.byref(i)), "tfs_matrix_get_info")
        return dict(n=i.n, nnz=i.nnz, levels_lower=i.levels_lower,
                    levels_upper=i.levels_upper, row_blocks=i.row_blocks,
                    device_bytes=i.device_bytes, inner_colours=i.inner_colours,
                    spmv_rows=i.spmv_rows, spmv_waves=i.spmv_waves,
                    inner_block=i.inner_block, inner_levels=i.inner_levels,
                    ell64_slices=i.ell64_slices, inner_order=i.inner_order,
                    sweep_groups=i.sweep_groups, inner_sweep_groups=i.inner_sweep_groups,
                    inner_brick_edge=i.inner_brick_edge)

    def tune_spmv(self):
        """Times the SpMV shapes and keeps the fastest; returns (rows, waves)."""
        r, w = ctypes.c_int(), ctypes.c_int()
        _check(self._lib.tfs_matrix_tune_spmv(self._h, ctypes.byref(r), ctypes.byref(w)),
               "tfs_matrix_tune_spmv")
        return r.value, w.value

    def set_spmv_shape(self, rows, waves):
        _check(self._lib.tfs_matrix_set_spmv_shape(self._h, int(rows), int(waves)),
               "tfs_matrix_set_spmv_shape")

    def set_sweep_schedule(self, schedule):
        """TFS_SWEEP_LEVELS / _STATIC / _DYNAMIC / _COOPERATIVE (otherwise timed on the
        first factor)."""
        _check(self._lib.tfs_matrix_set_sweep_schedule(self._h, int(schedule)),
               "tfs_matrix_set_sweep_schedule")

    def set_refinement(self, outer):
        """TFS_REFINE_RELIABLE (default), TFS_REFINE_FCG or TFS_REFINE_DEFECT for
        solve_mixed."""
        _check(self._lib.tfs_matrix_set_refinement(self._h, int(outer)),
               "tfs_matrix_set_refinement")

    def set_inner_block(self, rows):
        """Rows per colour block of the inner order (1 = plain colouring)."""
        _check(self._lib.tfs_matrix_set_inner_block(self._h, int(rows)),
               "tfs_matrix_set_inner_block")

    def set_inner_order(self, order):
        """TFS_INNER_REFERENCE, TFS_INNER_COLOURED, TFS_INNER_BRICKS or
        TFS_INNER_AUTO (default) for solve_mixed."""
        _check(self._lib.tfs_matrix_set_inner_order(self._h, int(order)),
               "tfs_matrix_set_inner_order")

    def set_values(self, val):
        if _is_device_tensor(val):
            _sync()
            rc = self._lib.tfs_matrix_set_values_device(self._h, _dptr(val, self.nnz, "val"))
        else:
            val = np.ascontiguousarray(val, dtype=np.float64)
            if val.size < self.nnz:
                raise ValueError("val shorter than nnz")
            rc = self._lib.tfs_matrix_set_values(self._h, val)
        _check(rc, "tfs_matrix_set_values")

    def get_values(self):
        """The caller-order values (host copy)."""
        v = np.empty(self.nnz, np.float64)
        _check(self._lib.tfs_matrix_get_values(self._h, v), "tfs_matrix_get_values")
        return v

    def values_changed(self):
        _check(self._lib.tfs_matrix_values_changed(self._h), "tfs_matrix_values_changed")

    def values_device_ptr(self):
        return self._lib.tfs_matrix_values_device(self._h)

    # ---- building blocks (torch GPU tensors) ----
    def spmv(self, x, y):
        _sync()
        _check(self._lib.tfs_spmv_device(self._h, _dptr(x, self.n, "x"), _dptr(y, self.n, "y")),
               "tfs_spmv_device")
        return y

    def residual(self, x, r):
        _sync()
        _check(self._lib.tfs_residual_device(self._h, _dptr(x, self.n, "x"),
                                             _dptr(r, self.n, "r")), "tfs_residual_device")
        return r

    def prec_form(self, prec):
        _check(self._lib.tfs_prec_form(self._h, prec), "tfs_prec_form")

    def prec_solve(self, prec, b, x):
        _sync()
        _check(self._lib.tfs_prec_solve_device(self._h, prec, _dptr(b, self.n, "b"),
                                               _dptr(x, self.n, "x")), "tfs_prec_solve_device")
        return x

    def prec_diag(self, out):
        _sync()
        _check(self._lib.tfs_prec_diag_device(self._h, _dptr(out, self.n, "out")),
               "tfs_prec_diag_device")
        return out

    def prec_diag32(self, prec, out):
        """The FP32 factor of the refinement solver, widened to FP64."""
        _sync()
        _check(self._lib.tfs_prec_diag32_device(self._h, prec, _dptr(out, self.n, "out")),
               "tfs_prec_diag32_device")
        return out

    def sweep_time(self, reps=20):
        """ms per FP64 incomplete-Cholesky sweep launch, back to back."""
        ms = ctypes.c_double()
        _check(self._lib.tfs_sweep_time(self._h, int(reps), ctypes.byref(ms)), "tfs_sweep_time")
        return ms.value

    def spmv_time_cold(self, x, y, reps=10):
        """ms per FP64 SpMV launch with the memory-side cache flushed before each."""
        _dptr(x, self.n, "x")
        _dptr(y, self.n, "y")
        ms = ctypes.c_double()
        _sync()
        _check(self._lib.tfs_spmv_time_cold(self._h, x.data_ptr(), y.data_ptr(), int(reps),
                                            ctypes.byref(ms)), "tfs_spmv_time_cold")
        return ms.value

    def spmv_time(self, x, y, reps=20):
        _sync()
        ms = ctypes.c_double()
        _check(self._lib.tfs_spmv_time(self._h, _dptr(x, self.n, "x"), _dptr(y, self.n, "y"),
                                       reps, ctypes.byref(ms)), "tfs_spmv_time")
        return ms.value


def _solve(method, A, x, r1, prec, niter, tol, ini_res=0.0):
    lib = A._lib
    it = ctypes.c_int(int(niter))
    ir = ctypes.c_double(float(ini_res))
    fr = ctypes.c_double(0.0)
    if _is_device_tensor(x) or _is_device_tensor(r1):
        _sync()
        rc = lib.tfs_solve_device(A.handle, method, prec, _dptr(x, A.n, "x"),
                                  _dptr(r1, A.n, "r1"), ctypes.byref(it), float(tol),
                                  ctypes.byref(ir), ctypes.byref(fr))
    else:
        for a, nm in ((x, "x"), (r1, "r1")):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64
                    and a.flags.c_contiguous and a.size >= A.n):
                raise ValueError(f"{nm}: expected a contiguous float64 numpy array")
        rc = lib.tfs_solve(A.handle, method, prec, x, r1, ctypes.byref(it), float(tol),
                           ctypes.byref(ir), ctypes.byref(fr))
    A.status = _check(rc, ("Cg", "Bicg", "Cgs")[method])
    return it.value, ir.value, fr.value


def Cg(A, x, r1, prec, niter, tol, ini_res=0.0):
    """Preconditioned CG, Sources/Process/Solvers_Mod/Cg.f90:2-183."""
    return _solve(CG, A, x, r1, prec, niter, tol, ini_res)


def Bicg(A, x, r1, prec, niter, tol, ini_res=0.0):
    """'BiCG' of Sources/Process/Solvers_Mod/Bicg.f90:2-196 (equal to Cg bit for bit)."""
    return _solve(BICG, A, x, r1, prec, niter, tol, ini_res)


def Cgs(A, x, r1, prec, niter, tol, ini_res=0.0):
    """Preconditioned CGS, Sources/Process/Solvers_Mod/Cgs.f90:2-232."""
    return _solve(CGS, A, x, r1, prec, niter, tol, ini_res)


def solve(method, A, x, r1, prec, niter, tol, ini_res=0.0):
    return _solve(_METHODS.get(method, method), A, x, r1, prec, niter, tol, ini_res)


def solve_mixed(A, x, r1, prec, niter, tol, inner_rtol=0.0, x_rtol=0.0):
    """FP32-inner PCG inside FP64 iterative refinement (GPU tensors only).

    Returns (steps, inner_iterations, ini_res, fin_res); fin_res is the true
    FP64 residual ||b - A x||, which is also left in r1.  inner_rtol: the
    FP32 recurrence's residual drop per residual replacement (reliable
    updates, default 0.1) or each inner solve's tolerance (flexible CG and
    defect correction, default 1e-2); 0 picks the default.  x_rtol > 0 (the
    flexible-CG and defect-correction loops) also requires the last
    correction to satisfy ||e|| <= x_rtol ||x||."""
    it = ctypes.c_int(int(niter))
    inner = ctypes.c_int(0)
    ir = ctypes.c_double(0.0)
    fr = ctypes.c_double(0.0)
    _sync()
    rc = A._lib.tfs_solve_mixed_device(A.handle, prec, _dptr(x, A.n, "x"), _dptr(r1, A.n, "r1"),
                                       ctypes.byref(it), ctypes.byref(inner), float(tol),
                                       float(inner_rtol), float(x_rtol), ctypes.byref(ir),
                                       ctypes.byref(fr))
    A.status = _check(rc, "solve_mixed")
    return it.value, inner.value, ir.value, fr.value


class KernelTimer:
    """HIP-event timing of one kernel class (TFS_KC_*) on the library stream."""
    SPMV, SWEEP, FACTOR = 0, 1, 2

    def __init__(self, kclass):
        self.kclass = kclass
        self.ms = 0.0
        self.launches = 0

    def __enter__(self):
        _check(_lib().tfs_timing_start(self.kclass), "tfs_timing_start")
        return self

    def __exit__(self, *exc):
        ms = ctypes.c_double()
        nl = ctypes.c_longlong()
        _lib().tfs_timing_stop(ctypes.byref(ms), ctypes.byref(nl))
        self.ms, self.launches = ms.value, nl.value
        return False


class Batch:
    """Many independent systems solved together (tfs_batch_*)."""

    def __init__(self, systems, base=0):
        """systems: list of (row, col, dia, val) in the reference CRS layout."""
        lib = _lib()
        if not getattr(lib, "_tfs_batch_typed", False):
            PP = ctypes.POINTER(ctypes.c_void_p)
            lib.tfs_batch_create.argtypes = [ctypes.c_int, _VP, _VP, PP, PP, PP, ctypes.c_int,
                                             ctypes.POINTER(_VP)]
            lib.tfs_batch_destroy.argtypes = [_VP]
            lib.tfs_batch_set_values.argtypes = [_VP, ctypes.c_int, _F64P]
            lib.tfs_batch_solve_device.argtypes = [_VP, ctypes.c_int, PP, PP, _VP, _VP, _VP, _VP,
                                                   _VP, _VP]
            lib.tfs_batch_solve_mixed_device.argtypes = [_VP, ctypes.c_int, PP, PP, ctypes.c_int,
                                                         _VP, ctypes.c_double, _VP, _VP, _VP,
                                                         _VP, _VP]
            lib.tfs_batch_get_info.argtypes = [_VP, _IP, _IP, ctypes.POINTER(ctypes.c_longlong),
                                               _IP, _IP, _IP]
            lib.tfs_batch_set_inner_order.argtypes = [_VP, ctypes.c_int]
            lib.tfs_batch_set_refinement.argtypes = [_VP, ctypes.c_int]
            lib.tfs_batch_set_inner_block.argtypes = [_VP, ctypes.c_int]
            lib._tfs_batch_typed = True
        self._lib = lib
        self._keep = []
        rows = [np.ascontiguousarray(s[0], np.int32) for s in systems]
        cols = [np.ascontiguousarray(s[1], np.int32) for s in systems]
        dias = [np.ascontiguousarray(s[2], np.int32) for s in systems]
        self._keep += rows + cols + dias
        self.nsys = len(systems)
        self.n = np.array([len(r) - 1 for r in rows], np.int32)
        self.nnz = np.array([int(r[-1]) - base for r in rows], np.int64)

        def ptrs(arrs):
            return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        h = _VP()
        _check(lib.tfs_batch_create(self.nsys, self.n.ctypes.data, self.nnz.ctypes.data,
                                    ptrs(rows), ptrs(cols), ptrs(dias), base, ctypes.byref(h)),
               "tfs_batch_create")
        self._h = h
        _track(self)
        for s, sysd in enumerate(systems):
            if len(sysd) > 3 and sysd[3] is not None:
                self.set_values(s, sysd[3])

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tfs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_values(self, s, val):
        val = np.ascontiguousarray(val, np.float64)
        _check(self._lib.tfs_batch_set_values(self._h, s, val), "tfs_batch_set_values")

    def set_inner_order(self, order):
        _check(self._lib.tfs_batch_set_inner_order(self._h, int(order)),
               "tfs_batch_set_inner_order")

    def set_refinement(self, outer):
        _check(self._lib.tfs_batch_set_refinement(self._h, int(outer)),
               "tfs_batch_set_refinement")

    def set_inner_block(self, rows):
        _check(self._lib.tfs_batch_set_inner_block(self._h, int(rows)),
               "tfs_batch_set_inner_block")

    def info(self):
        ns, nt, gl, sr, sw = (ctypes.c_int() for _ in range(5))
        nz = ctypes.c_longlong()
        self._lib.tfs_batch_get_info(self._h, ctypes.byref(ns), ctypes.byref(nt),
                                     ctypes.byref(nz), ctypes.byref(gl), ctypes.byref(sr),
                                     ctypes.byref(sw))
        return dict(nsys=ns.value, n_total=nt.value, nnz_total=nz.value,
                    graph_launches=gl.value, spmv_rows=sr.value, spmv_waves=sw.value)

    def _vecs(self, xs, rs):
        for k, (x, r) in enumerate(zip(xs, rs)):
            _dptr(x, int(self.n[k]), "x")
            _dptr(r, int(self.n[k]), "r1")
        px = (ctypes.c_void_p * self.nsys)(*[x.data_ptr() for x in xs])
        pr = (ctypes.c_void_p * self.nsys)(*[r.data_ptr() for r in rs])
        return px, pr

    def solve(self, xs, rs, prec, caps, tols):
        """Cg per system; returns arrays (niter, ini_res, fin_res, status)."""
        px, pr = self._vecs(xs, rs)
        caps = np.ascontiguousarray(caps, np.int32)
        tols = np.ascontiguousarray(tols, np.float64)
        niter = np.zeros(self.nsys, np.int32)
        ini = np.zeros(self.nsys)
        fin = np.zeros(self.nsys)
        st = np.zeros(self.nsys, np.int32)
        _sync()
        _check(self._lib.tfs_batch_solve_device(self._h, prec, px, pr, caps.ctypes.data,
                                                tols.ctypes.data, niter.ctypes.data,
                                                ini.ctypes.data, fin.ctypes.data, st.ctypes.data),
               "tfs_batch_solve_device")
        return niter, ini, fin, st

    def solve_mixed(self, xs, rs, prec, cap, tols, inner_rtol=0.0):
        """FP32-inner refinement per system; arrays (steps, inner, ini, fin, status)."""
        px, pr = self._vecs(xs, rs)
        tols = np.ascontiguousarray(tols, np.float64)
        steps = np.zeros(self.nsys, np.int32)
        inner = np.zeros(self.nsys, np.int32)
        ini = np.zeros(self.nsys)
        fin = np.zeros(self.nsys)
        st = np.zeros(self.nsys, np.int32)
        _sync()
        _check(self._lib.tfs_batch_solve_mixed_device(self._h, prec, px, pr, int(cap),
                                                      tols.ctypes.data, float(inner_rtol),
                                                      steps.ctypes.data, inner.ctypes.data,
                                                      ini.ctypes.data, fin.ctypes.data,
                                                      st.ctypes.data), "tfs_batch_solve_mixed")
        return steps, inner, ini, fin, st


class Server:
    """Continuous-batching request server (tfs_server_*): host numpy in/out."""

    def __init__(self, hbm_budget_bytes, workers=4):
        lib = _lib()
        if not getattr(lib, "_tfs_server_typed", False):
            LLP = ctypes.POINTER(ctypes.c_longlong)
            lib.tfs_server_create.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(_VP)]
            lib.tfs_server_destroy.argtypes = [_VP]
            lib.tfs_server_add_case.argtypes = [_VP, ctypes.c_int, ctypes.c_longlong, _I32P, _I32P,
                                                _I32P, _F64P, ctypes.c_int, _IP]
            lib.tfs_server_submit.argtypes = [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _F64P,
                                              ctypes.c_double, ctypes.c_int, ctypes.c_int, LLP]
            lib.tfs_server_wait.argtypes = [_VP, ctypes.c_longlong, _VP, _IP, _DP, _DP, _IP, _DP,
                                            _DP, _DP]
            lib.tfs_server_stats.argtypes = [_VP, LLP, LLP, _IP, LLP]
            lib.tfs_server_cancel.argtypes = [_VP, ctypes.c_longlong]
            lib._tfs_server_typed = True
        self._lib = lib
        self._n = {}
        h = _VP()
        _check(lib.tfs_server_create(int(hbm_budget_bytes), int(workers), ctypes.byref(h)),
               "tfs_server_create")
        self._h = h
        _track(self)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tfs_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_case(self, row, col, dia, val, base=0):
        row = np.ascontiguousarray(row, np.int32)
        n = len(row) - 1
        cid = ctypes.c_int()
        _check(self._lib.tfs_server_add_case(self._h, n, int(row[-1]) - base, row,
                                             np.ascontiguousarray(col, np.int32),
                                             np.ascontiguousarray(dia, np.int32),
                                             np.ascontiguousarray(val, np.float64), base,
                                             ctypes.byref(cid)), "tfs_server_add_case")
        self._n[cid.value] = n
        return cid.value

    def submit(self, case_id, method, prec, b, tol, cap, release=False):
        req = ctypes.c_longlong()
        _check(self._lib.tfs_server_submit(self._h, case_id, _METHODS.get(method, method), prec,
                                           np.ascontiguousarray(b, np.float64), float(tol),
                                           int(cap), int(release), ctypes.byref(req)),
               "tfs_server_submit")
        return (req.value, self._n[case_id])

    def wait(self, handle, want_x=True):
        req, n = handle
        x = np.zeros(n) if want_x else None
        it, st = ctypes.c_int(), ctypes.c_int()
        ir, fr, t0, t1, t2 = (ctypes.c_double() for _ in range(5))
        _check(self._lib.tfs_server_wait(self._h, req, x.ctypes.data if want_x else None,
                                         ctypes.byref(it), ctypes.byref(ir), ctypes.byref(fr),
                                         ctypes.byref(st), ctypes.byref(t0), ctypes.byref(t1),
                                         ctypes.byref(t2)), "tfs_server_wait")
        return dict(x=x, niter=it.value, ini_res=ir.value, fin_res=fr.value, status=st.value,
                    t_submit=t0.value, t_start=t1.value, t_end=t2.value)

    def cancel(self, handle):
        """True if the request was withdrawn before it started."""
        return self._lib.tfs_server_cancel(self._h, handle[0]) == 0

    def stats(self):
        a, b, d = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        c = ctypes.c_int()
        self._lib.tfs_server_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                   ctypes.byref(d))
        return dict(hbm_in_use=a.value, hbm_peak=b.value, resident=c.value, completed=d.value)


class Assembly:
    """Device-side face-loop assembly of a Matrix's values (tfs_assembly_*):
    faces_c / pos / dia as the topology builds them (1-based by default);
    coefficients are GPU tensors (a12, a21 per face, float64; bnd_on per face,
    uint8; cell, sav per cell, float64)."""

    def __init__(self, A, faces_c, pos, dia, base=1):
        lib = A._lib
        if not getattr(lib, "_tfs_asm_typed", False):
            lib.tfs_assembly_create.argtypes = [_VP, ctypes.c_longlong, _I32P, _I32P, _I32P,
                                                ctypes.c_int, ctypes.POINTER(_VP)]
            lib.tfs_assembly_destroy.argtypes = [_VP]
            lib.tfs_assemble_device.argtypes = [_VP, _VP, _VP, _VP, _VP, ctypes.c_double, _VP]
            lib._tfs_asm_typed = True
        self._lib = lib
        self.A = A
        faces_c = np.ascontiguousarray(faces_c, np.int32)
        self.nf = faces_c.size // 2
        h = _VP()
        _check(lib.tfs_assembly_create(A.handle, self.nf, faces_c,
                                       np.ascontiguousarray(pos, np.int32),
                                       np.ascontiguousarray(dia, np.int32), base, ctypes.byref(h)),
               "tfs_assembly_create")
        self._h = h
        _track(self)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tfs_assembly_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def assemble(self, a12, a21=None, bnd_on=None, cell=None, urf=1.0, sav=None):
        def ptr(t, n, what):
            if t is None:
                return None
            if not _is_device_tensor(t) or t.numel() != n or not t.is_contiguous():
                raise ValueError(f"{what}: contiguous GPU tensor of {n} elements expected")
            return t.data_ptr()
        n = self.A.n
        _sync()
        _check(self._lib.tfs_assemble_device(self._h, ptr(a12, self.nf, "a12"),
                                             ptr(a21, self.nf, "a21"),
                                             ptr(bnd_on, self.nf, "bnd_on"),
                                             ptr(cell, n, "cell"), float(urf),
                                             ptr(sav, n, "sav")), "tfs_assemble_device")


def topology_device(n_cells, faces_c, base=1, with_pos=True):
    """Matrix_Mod_Topology on the GPU (tfs_topology_build_device): faces_c is
    a contiguous int32 GPU tensor of 2 n_faces entries (c1, c2 per face,
    1-based, c2 <= 0 boundary).  Returns GPU int32 tensors (row, col, dia,
    pos) in index base `base` (pos None unless with_pos)."""
    import torch
    lib = _lib()
    if not getattr(lib, "_tfs_topo_typed", False):
        lib.tfs_topology_build_device.argtypes = [ctypes.c_int, ctypes.c_longlong, _VP,
                                                  ctypes.c_int, ctypes.c_longlong, _VP, _VP,
                                                  _VP, _VP, ctypes.POINTER(ctypes.c_longlong)]
        lib._tfs_topo_typed = True
    if not (_is_device_tensor(faces_c) and faces_c.dtype == torch.int32
            and faces_c.is_contiguous() and faces_c.numel() % 2 == 0):
        raise ValueError("faces_c: contiguous int32 GPU tensor of 2 n_faces entries expected")
    nf = faces_c.numel() // 2
    cap = int(n_cells) + 2 * nf
    opts = dict(dtype=torch.int32, device=faces_c.device)
    row = torch.empty(int(n_cells) + 1, **opts)
    col = torch.empty(max(cap, 1), **opts)
    dia = torch.empty(int(n_cells), **opts)
    pos = torch.empty(max(2 * nf, 2), **opts) if with_pos else None
    nnz = ctypes.c_longlong(0)
    _sync()
    _check(lib.tfs_topology_build_device(int(n_cells), nf, faces_c.data_ptr() if nf else None,
                                         int(base), cap, row.data_ptr(), col.data_ptr(),
                                         dia.data_ptr(), pos.data_ptr() if with_pos else None,
                                         ctypes.byref(nnz)), "tfs_topology_build_device")
    return row, col[:nnz.value], dia, (pos[:2 * nf] if with_pos else None)


TFS_BC_WALL, TFS_BC_SYMMETRY, TFS_BC_PRESSURE, TFS_BC_BUFFER = 0, 1, 2, 3


class Gradient:
    """Least-squares cell gradients on the GPU (tfs_gradient_*): the
    counterparts of Compute_Gradient_Matrix, GraPhi and GradP.  faces_c, d
    (3, n_faces), xc (3, n_cells + n_bnd) and bc (kind per boundary cell) are
    host arrays as the mesh readers return them; phi and the gradient
    components are GPU float64 tensors."""

    def __init__(self, n_cells, n_bnd_cells, faces_c, d, xc=None, bc=None):
        lib = _lib()
        if not getattr(lib, "_tfs_grad_typed", False):
            lib.tfs_gradient_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_longlong,
                                                _I32P, _F64P, _VP, _VP, ctypes.POINTER(_VP)]
            lib.tfs_gradient_destroy.argtypes = [_VP]
            lib.tfs_gradient_matrix_get.argtypes = [_VP, _F64P]
            lib.tfs_grad_phi_device.argtypes = [_VP] + [_VP] * 4
            lib.tfs_grad_p_device.argtypes = [_VP] + [_VP] * 4
            lib.tfs_correct_velocity_device.argtypes = [_VP] + [_VP] * 14 + \
                [ctypes.c_double, ctypes.c_double, _VP, _VP, _DP, _DP, _DP]
            lib.tfs_pressure_fractional_device.argtypes = [_VP] * 16
            lib.tfs_pressure_simple_device.argtypes = [_VP] * 17
            lib._tfs_grad_typed = True
        self._lib = lib
        self.n, self.nb = int(n_cells), int(n_bnd_cells)
        faces_c = np.ascontiguousarray(faces_c, np.int32)
        self.nf = faces_c.size // 2  # the mesh's face count: every per-face array has it
        d = np.ascontiguousarray(d, np.float64).ravel()
        self._keep = [None if xc is None else np.ascontiguousarray(xc, np.float64).ravel(),
                      None if bc is None else np.ascontiguousarray(bc, np.uint8)]
        h = _VP()
        _check(lib.tfs_gradient_create(self.n, self.nb, faces_c.size // 2, faces_c, d,
                                       None if xc is None else self._keep[0].ctypes.data,
                                       None if bc is None else self._keep[1].ctypes.data,
                                       ctypes.byref(h)), "tfs_gradient_create")
        self._h = h
        _track(self)

    def matrix(self):
        """G(1:6, c) as an (n_cells, 6) host array."""
        g = np.zeros((self.n, 6))
        _check(self._lib.tfs_gradient_matrix_get(self._h, g), "tfs_gradient_matrix_get")
        return g

    def _args(self, phi, px, py, pz):
        return (_dptr(phi, self.n + self.nb, "phi"), _dptr(px, self.n, "px"),
                _dptr(py, self.n, "py"), _dptr(pz, self.n, "pz"))

    def grad_phi(self, phi, px, py, pz):
        _sync()
        _check(self._lib.tfs_grad_phi_device(self._h, *self._args(phi, px, py, pz)),
               "tfs_grad_phi_device")

    def grad_p(self, phi, px, py, pz):
        _sync()
        _check(self._lib.tfs_grad_p_device(self._h, *self._args(phi, px, py, pz)),
               "tfs_grad_p_device")

    def correct_velocity(self, vol, sav, px, py, pz, u, v, w, pp, coef, den, scoef, dt, visc,
                         flux, b, bou=None, in_m=None):
        """Correct_Velocity on the GPU: u, v, w (cells and boundary cells) and
        flux (faces) are updated in place, b (cells) receives the mass error;
        returns (errmax, cfl_max, pe_max).  coef per face = A % val(A % pos(1, s));
        den per cell = DENc(material(c)); in_m (uint8 per face) marks the
        faces of the material whose CFL / Peclet numbers are reported."""
        n, na, nf = self.n, self.n + self.nb, self.nf
        for t, what in ((coef, "coef"), (scoef, "scoef"), (flux, "flux"), (in_m, "in_m")):
            _check_count(t, nf, what)
        for t, m, what in ((u, na, "u"), (v, na, "v"), (w, na, "w"), (pp, na, "pp"),
                           (vol, n, "vol"), (sav, n, "sav"), (px, n, "px"), (py, n, "py"),
                           (pz, n, "pz"), (den, n, "den"), (b, n, "b"), (bou, self.nb, "bou")):
            _check_count(t, m, what, exact=False)

        def opt(t, m, what, dtype=None):
            if t is None:
                return None
            import torch
            want = torch.float64 if dtype is None else dtype
            if not (_is_device_tensor(t) and t.dtype == want and t.is_contiguous()
                    and t.numel() >= m):
                raise ValueError(f"{what}: contiguous GPU tensor of {m} entries expected")
            return t.data_ptr()
        import torch
        em, cm, pm = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _sync()
        _check(self._lib.tfs_correct_velocity_device(
            self._h, _dptr(vol, n, "vol"), _dptr(sav, n, "sav"), _dptr(px, n, "px"),
            _dptr(py, n, "py"), _dptr(pz, n, "pz"), _dptr(u, na, "u"), _dptr(v, na, "v"),
            _dptr(w, na, "w"), _dptr(pp, na, "pp"), _dptr(coef, nf, "coef"),
            opt(bou, self.nb, "bou"), _dptr(den, n, "den"), opt(in_m, nf, "in_m", torch.uint8),
            _dptr(scoef, nf, "scoef"), float(dt), float(visc), _dptr(flux, nf, "flux"),
            _dptr(b, n, "b"), ctypes.byref(em), ctypes.byref(cm), ctypes.byref(pm)),
            "tfs_correct_velocity_device")
        return em.value, cm.value, pm.value

    def pressure_fractional(self, f, s, vol, sav, u, v, w, px, py, pz, p, den, flux, a12, b):
        """Compute_Pressure_Fractional's face loop on the GPU: flux and a12 per
        face, b per cell (GPU float64 tensors; s holds sx, sy, sz blocks)."""
        n, na, nf = self.n, self.n + self.nb, self.nf
        for t, what in ((f, "f"), (flux, "flux"), (a12, "a12")):
            _check_count(t, nf, what)
        _check_count(s, 3 * nf, "s (sx, sy, sz blocks)")
        for t, m, what in ((u, na, "u"), (v, na, "v"), (w, na, "w"), (p, na, "p"),
                           (den, na, "den"), (vol, n, "vol"), (sav, n, "sav"), (px, n, "px"),
                           (py, n, "py"), (pz, n, "pz"), (b, n, "b")):
            _check_count(t, m, what, exact=False)
        _sync()
        _check(self._lib.tfs_pressure_fractional_device(
            self._h, _dptr(f, nf, "f"), _dptr(s, 3 * nf, "s"), _dptr(vol, n, "vol"),
            _dptr(sav, n, "sav"), _dptr(u, na, "u"), _dptr(v, na, "v"), _dptr(w, na, "w"),
            _dptr(px, n, "px"), _dptr(py, n, "py"), _dptr(pz, n, "pz"), _dptr(p, na, "p"),
            _dptr(den, na, "den"), _dptr(flux, nf, "flux"), _dptr(a12, nf, "a12"),
            _dptr(b, n, "b")), "tfs_pressure_fractional_device")

    def pressure_simple(self, f, s, vol, sav, u, v, w, px, py, pz, p, den, flux, a12, bnd_on, b):
        """Compute_Pressure_Simple's face loop on the GPU: flux, a12 and bnd_on
        (uint8) per face, b per cell (GPU tensors; s holds sx, sy, sz blocks);
        the matrix then comes from Assembly.assemble(a12, bnd_on=bnd_on)."""
        import torch
        n, na, nf = self.n, self.n + self.nb, self.nf
        for t, what in ((f, "f"), (flux, "flux"), (a12, "a12"), (bnd_on, "bnd_on")):
            _check_count(t, nf, what)
        _check_count(s, 3 * nf, "s (sx, sy, sz blocks)")
        for t, m, what in ((u, na, "u"), (v, na, "v"), (w, na, "w"), (p, na, "p"),
                           (den, na, "den"), (vol, n, "vol"), (sav, n, "sav"), (px, n, "px"),
                           (py, n, "py"), (pz, n, "pz"), (b, n, "b")):
            _check_count(t, m, what, exact=False)
        if not (_is_device_tensor(bnd_on) and bnd_on.dtype == torch.uint8
                and bnd_on.is_contiguous()):
            raise ValueError("bnd_on: contiguous uint8 GPU tensor expected")
        _sync()
        _check(self._lib.tfs_pressure_simple_device(
            self._h, _dptr(f, nf, "f"), _dptr(s, 3 * nf, "s"), _dptr(vol, n, "vol"),
            _dptr(sav, n, "sav"), _dptr(u, na, "u"), _dptr(v, na, "v"), _dptr(w, na, "w"),
            _dptr(px, n, "px"), _dptr(py, n, "py"), _dptr(pz, n, "pz"), _dptr(p, na, "p"),
            _dptr(den, na, "den"), _dptr(flux, nf, "flux"), _dptr(a12, nf, "a12"),
            bnd_on.data_ptr(), _dptr(b, n, "b")), "tfs_pressure_simple_device")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tfs_gradient_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
