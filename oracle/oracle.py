"""ctypes access to the oracle (TEST INFRASTRUCTURE ONLY).

* ``Oracle``    -- the C restatement, oracle/_build/libtfs_oracle.so.
* ``Reference`` -- the reference's own Fortran solver path compiled with the
  flags of its FAST build, oracle/_ref/libtfs_ref.so (only where it was built).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline.  Matrices are
passed 0-based (row/col/dia as numpy int32, val float64); ``Reference`` shifts
them to the reference's 1-based numbering.
"""
import ctypes
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_ORACLE = HERE / "_build" / "libtfs_oracle.so"
LIB_REF = HERE / "_ref" / "libtfs_ref.so"

METHODS = {"cg": 0, "bicg": 1, "cgs": 2}

_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class _OrMatrix(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int),
                ("row", ctypes.c_void_p), ("col", ctypes.c_void_p),
                ("dia", ctypes.c_void_p), ("val", ctypes.c_void_p)]


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """The C restatement of Solvers_Mod / Matrix_Mod_Topology."""

    def __init__(self, path=LIB_ORACLE):
        if not pathlib.Path(path).exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        lib = ctypes.CDLL(str(path))
        mp = ctypes.POINTER(_OrMatrix)
        lib.or_prec_form.argtypes = [mp, ctypes.c_int, _F64P]
        lib.or_prec_solve.argtypes = [mp, _F64P, ctypes.c_int, _F64P, _F64P]
        lib.or_residual.argtypes = [mp, _F64P, _F64P]
        lib.or_spmv.argtypes = [mp, _F64P, _F64P]
        for f in (lib.or_cg, lib.or_bicg, lib.or_cgs):
            f.restype = ctypes.c_int
            f.argtypes = [mp, _F64P, _F64P, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                          ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double)]
        lib.or_topology_nnz.restype = ctypes.c_int
        lib.or_topology_nnz.argtypes = [ctypes.c_int, ctypes.c_int, _I32P]
        lib.or_topology.argtypes = [ctypes.c_int, ctypes.c_int, _I32P] + [_I32P] * 5
        lib.or_assemble.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, _I32P, ctypes.c_int,
                                    _F64P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_double, _F64P, ctypes.c_void_p]
        _VP = ctypes.c_void_p
        lib.or_gradient_matrix.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _F64P, _F64P]
        lib.or_graphi.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _F64P, _F64P, _F64P,
                                  ctypes.c_int, _VP, _F64P]
        lib.or_gradp.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _F64P, _F64P, _F64P,
                                 ctypes.c_int, _VP, _F64P, _F64P, _F64P, _F64P]
        lib.or_correct_velocity.restype = ctypes.c_double
        lib.or_correct_velocity.argtypes = [ctypes.c_int, ctypes.c_int, _I32P] + [_VP] * 16 + \
            [ctypes.c_double, ctypes.c_double, _VP, _VP, _VP, _VP]
        lib.or_pressure_fractional.argtypes = [ctypes.c_int, ctypes.c_int, _I32P] + [_VP] * 15
        lib.or_pressure_simple.argtypes = [ctypes.c_int, ctypes.c_int, _I32P] + [_VP] * 14 + \
            [_I32P, _I32P, ctypes.c_int] + [_VP] * 5
        self.lib = lib

    def pressure_simple(self, n_cells, faces_c, f, s, d, vol, sav, u, v, w, p, px, py, pz, den,
                        bc, pos, dia, nnz):
        """Compute_Pressure_Simple's face loop: dict(flux, a12, bnd_on, val, b);
        faces_c, pos, dia 1-based; bc = kind per boundary cell (or None: walls)."""
        faces_c, pos, dia = _c32(faces_c), _c32(pos), _c32(dia)
        nf = faces_c.size // 2
        keep = [_c64(a).ravel() for a in (f, s, d, vol, sav, u, v, w, p, px, py, pz, den)]
        bck = None if bc is None else np.ascontiguousarray(bc, np.uint8)
        flux, a12, b = np.zeros(nf), np.zeros(nf), np.zeros(n_cells)
        bnd_on, val = np.zeros(nf, np.uint8), np.zeros(nnz)
        self.lib.or_pressure_simple(n_cells, nf, faces_c, *(a.ctypes.data for a in keep),
                                    None if bck is None else bck.ctypes.data, pos, dia, int(nnz),
                                    flux.ctypes.data, a12.ctypes.data, bnd_on.ctypes.data,
                                    val.ctypes.data, b.ctypes.data)
        return dict(flux=flux, a12=a12, bnd_on=bnd_on, val=val, b=b)

    def pressure_fractional(self, n_cells, faces_c, f, s, vol, sav, u, v, w, px, py, pz, p, den):
        """Compute_Pressure_Fractional's face loop: (flux, a12, b)."""
        faces_c = _c32(faces_c)
        nf = faces_c.size // 2
        keep = [_c64(a).ravel() for a in (f, s, vol, sav, u, v, w, px, py, pz, p, den)]
        flux, a12, b = np.zeros(nf), np.zeros(nf), np.zeros(n_cells)
        self.lib.or_pressure_fractional(n_cells, nf, faces_c, *(a.ctypes.data for a in keep),
                                        flux.ctypes.data, a12.ctypes.data, b.ctypes.data)
        return flux, a12, b

    def correct_velocity(self, n_cells, faces_c, vol, sav, px, py, pz, u, v, w, pp, coef, scoef,
                         d, dt, visc, flux, den, bc=None, bou=None, in_m=None):
        """Correct_Velocity: returns dict(u, v, w, flux, b, errmax, cfl, pe);
        the inputs are not modified."""
        faces_c = _c32(faces_c)
        nf = faces_c.size // 2
        u, v, w, flux = (_c64(a).copy() for a in (u, v, w, flux))
        keep = [_c64(a) for a in (vol, sav, px, py, pz, pp, coef, den, scoef)] + [_c64(d).ravel()]
        bcb = None if bc is None else np.ascontiguousarray(bc, np.uint8)
        bo = None if bou is None else _c64(bou)
        im = None if in_m is None else np.ascontiguousarray(in_m, np.uint8)
        b = np.zeros(n_cells)
        cfl, pe = ctypes.c_double(), ctypes.c_double()
        ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        vol_, sav_, px_, py_, pz_, pp_, coef_, den_, scoef_, d_ = keep
        err = self.lib.or_correct_velocity(
            n_cells, nf, faces_c, ptr(vol_), ptr(sav_), ptr(px_), ptr(py_), ptr(pz_), ptr(u),
            ptr(v), ptr(w), ptr(pp_), ptr(coef_), ptr(bo), ptr(bcb), ptr(den_), ptr(im),
            ptr(scoef_), ptr(d_), float(dt), float(visc), ptr(flux), ptr(b),
            ctypes.addressof(cfl), ctypes.addressof(pe))
        return dict(u=u, v=v, w=w, flux=flux, b=b, errmax=err, cfl=cfl.value, pe=pe.value)

    # ---- least-squares gradients: d (3, n_faces); g (n_cells, 6) = G(1:6, c)
    def gradient_matrix(self, n_cells, faces_c, d):
        faces_c = _c32(faces_c)
        g = np.zeros((n_cells, 6))
        self.lib.or_gradient_matrix(n_cells, faces_c.size // 2, faces_c, _c64(d).ravel(), g)
        return g

    def graphi(self, n_cells, faces_c, d, g, phi, bc=None):
        """(phi_x, phi_y, phi_z): GraPhi for i = 1, 2, 3 (Boundary = .TRUE.)."""
        faces_c = _c32(faces_c)
        bcb = None if bc is None else np.ascontiguousarray(bc, np.uint8)
        out = []
        for i in (1, 2, 3):
            o = np.zeros(n_cells)
            self.lib.or_graphi(n_cells, faces_c.size // 2, faces_c, _c64(d).ravel(), _c64(g), _c64(phi),
                               i, None if bcb is None else bcb.ctypes.data, o)
            out.append(o)
        return tuple(out)

    def gradp(self, n_cells, faces_c, d, g, xc, phi, bc=None):
        """GradP: (phi with its boundary part rewritten, px, py, pz)."""
        faces_c = _c32(faces_c)
        xc = _c64(xc)
        n_bnd = xc.size // 3 - n_cells
        bcb = None if bc is None else np.ascontiguousarray(bc, np.uint8)
        phi = _c64(phi).copy()
        px, py, pz = (np.zeros(n_cells) for _ in range(3))
        self.lib.or_gradp(n_cells, faces_c.size // 2, faces_c, _c64(d).ravel(), _c64(g), xc.ravel(), n_bnd,
                          None if bcb is None else bcb.ctypes.data, phi, px, py, pz)
        return phi, px, py, pz

    def _mat(self, row, col, dia, val):
        keep = (_c32(row), _c32(col), _c32(dia), _c64(val))
        m = _OrMatrix(len(keep[0]) - 1, keep[0].ctypes.data, keep[1].ctypes.data,
                      keep[2].ctypes.data, keep[3].ctypes.data)
        return m, keep

    def prec_form(self, row, col, dia, val, prec):
        m, keep = self._mat(row, col, dia, val)
        d = np.zeros(m.n, np.float64)
        self.lib.or_prec_form(ctypes.byref(m), prec, d)
        return d

    def prec_solve(self, row, col, dia, val, d, prec, b):
        m, keep = self._mat(row, col, dia, val)
        x = np.zeros(m.n, np.float64)
        self.lib.or_prec_solve(ctypes.byref(m), _c64(d), prec, x, _c64(b))
        return x

    def residual(self, row, col, dia, val, x, r):
        m, keep = self._mat(row, col, dia, val)
        r = _c64(r).copy()
        self.lib.or_residual(ctypes.byref(m), _c64(x), r)
        return r

    def spmv(self, row, col, dia, val, x):
        m, keep = self._mat(row, col, dia, val)
        y = np.zeros(m.n, np.float64)
        self.lib.or_spmv(ctypes.byref(m), _c64(x), y)
        return y

    def solve(self, method, row, col, dia, val, b, prec=2, niter=1000, tol=1e-8, x0=None,
              ini_res=0.0):
        """Returns dict(x, r, niter, ini_res, fin_res) -- the reference outputs."""
        m, keep = self._mat(row, col, dia, val)
        x = np.zeros(m.n, np.float64) if x0 is None else _c64(x0).copy()
        r = _c64(b).copy()
        it = ctypes.c_int(niter)
        ir = ctypes.c_double(ini_res)
        fr = ctypes.c_double(0.0)
        fn = {0: self.lib.or_cg, 1: self.lib.or_bicg, 2: self.lib.or_cgs}[METHODS.get(method, method)]
        fn(ctypes.byref(m), x, r, prec, ctypes.byref(it), tol, ctypes.byref(ir), ctypes.byref(fr))
        return dict(x=x, r=r, niter=it.value, ini_res=ir.value, fin_res=fr.value)

    def assemble(self, n_cells, faces_c, pos, dia, nnz, a12, a21=None, bnd_on=None, cell=None,
                 urf=1.0, with_sav=False):
        """Face-loop assembly; faces_c/pos/dia 1-based.  Returns val (and sav)."""
        faces_c, pos, dia = _c32(faces_c), _c32(pos), _c32(dia)
        nf = faces_c.size // 2
        keep = [None if v is None else _c64(v) for v in (a21, cell)]
        bo = None if bnd_on is None else np.ascontiguousarray(bnd_on, np.uint8)
        val = np.zeros(nnz, np.float64)
        sav = np.zeros(n_cells, np.float64) if with_sav else None
        self.lib.or_assemble(n_cells, nf, faces_c, pos, dia, nnz, _c64(a12),
                             None if keep[0] is None else keep[0].ctypes.data,
                             None if bo is None else bo.ctypes.data,
                             None if keep[1] is None else keep[1].ctypes.data, float(urf), val,
                             None if sav is None else sav.ctypes.data)
        return (val, sav) if with_sav else val

    def topology(self, n_cells, faces_c):
        faces_c = _c32(faces_c)
        nf = faces_c.size // 2
        nnz = self.lib.or_topology_nnz(n_cells, nf, faces_c)
        row = np.zeros(n_cells + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        dia = np.zeros(n_cells, np.int32)
        mir = np.zeros(nnz, np.int32)
        pos = np.zeros(2 * nf, np.int32)
        self.lib.or_topology(n_cells, nf, faces_c, row, col, dia, mir, pos)
        return dict(row=row, col=col, dia=dia, mir=mir, pos=pos)


def reference_available(path=LIB_REF):
    return pathlib.Path(path).exists()


class Reference:
    """The reference solver path itself (compiled from its sources)."""

    def __init__(self, path=LIB_REF):
        if not pathlib.Path(path).exists():
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        lib = ctypes.CDLL(str(path))
        lib.tfs_ref_set_matrix.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _I32P, _I32P, _F64P]
        lib.tfs_ref_solve.restype = ctypes.c_double
        lib.tfs_ref_solve.argtypes = [ctypes.c_int, ctypes.c_int, _F64P, _F64P, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int), ctypes.c_double,
                                      ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
        lib.tfs_ref_prec.argtypes = [ctypes.c_int, ctypes.c_int, _F64P, _F64P, _F64P]
        lib.tfs_ref_residual.argtypes = [ctypes.c_int, _F64P, _F64P]
        lib.tfs_ref_topology.restype = ctypes.c_int
        lib.tfs_ref_topology.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _I32P,
                                         ctypes.c_int] + [_I32P] * 5
        lib.tfs_ref_pressure_matrix.restype = ctypes.c_int
        lib.tfs_ref_pressure_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _I32P,
                                                ctypes.c_double, _F64P, ctypes.c_int, _F64P]
        if hasattr(lib, "tfs_ref_mesh_load"):
            lib.tfs_ref_mesh_load.restype = ctypes.c_int
            lib.tfs_ref_mesh_load.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                                              _I32P]
            lib.tfs_ref_mesh_grad.restype = ctypes.c_int
            lib.tfs_ref_mesh_grad.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 6
            lib.tfs_ref_mesh_correct.restype = ctypes.c_int
            lib.tfs_ref_mesh_correct.argtypes = [ctypes.c_void_p] * 10 + \
                [ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
            lib.tfs_ref_mesh_press_frac.restype = ctypes.c_int
            lib.tfs_ref_mesh_press_frac.argtypes = [ctypes.c_void_p] * 8 + \
                [ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double] + \
                [ctypes.c_void_p] * 4
            lib.tfs_ref_mesh_press_simple.restype = ctypes.c_int
            lib.tfs_ref_mesh_press_simple.argtypes = [ctypes.c_void_p] * 8 + \
                [ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                 ctypes.c_double] + [ctypes.c_void_p] * 4
            lib.tfs_ref_mesh_get.restype = ctypes.c_int
            lib.tfs_ref_mesh_get.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                             ctypes.c_void_p]
        self.lib = lib
        self.n = 0

    def set_matrix(self, row, col, dia, val):
        """0-based CRS in; stored 1-based inside the reference Matrix_Type."""
        row = _c32(row) + 1
        col = _c32(col) + 1
        dia = _c32(dia) + 1
        self.n = len(row) - 1
        self.lib.tfs_ref_set_matrix(self.n, int(row[-1] - 1), row, col, dia, _c64(val))

    def solve(self, method, b, prec=2, niter=1000, tol=1e-8, x0=None, ini_res=0.0):
        """Returns dict(x, r, niter, ini_res, fin_res, seconds)."""
        x = np.zeros(self.n, np.float64) if x0 is None else _c64(x0).copy()
        r = _c64(b).copy()
        it = ctypes.c_int(niter)
        ir = ctypes.c_double(ini_res)
        fr = ctypes.c_double(0.0)
        t = self.lib.tfs_ref_solve(METHODS.get(method, method), self.n, x, r, prec,
                                   ctypes.byref(it), tol, ctypes.byref(ir), ctypes.byref(fr))
        return dict(x=x, r=r, niter=it.value, ini_res=ir.value, fin_res=fr.value, seconds=t)

    def prec(self, prec, b):
        z = np.zeros(self.n, np.float64)
        d = np.zeros(self.n, np.float64)
        self.lib.tfs_ref_prec(self.n, prec, _c64(b), z, d)
        return z, d

    def residual(self, x, r):
        r = _c64(r).copy()
        self.lib.tfs_ref_residual(self.n, _c64(x), r)
        return r

    def topology(self, n_cells, faces_c, n_bnd=None):
        faces_c = _c32(faces_c)
        nf = faces_c.size // 2
        if n_bnd is None:
            n_bnd = int(max(0, -faces_c[1::2].min())) if nf else 0
        cap = n_cells + 2 * nf
        row = np.zeros(n_cells + 1, np.int32)
        col = np.zeros(cap, np.int32)
        dia = np.zeros(n_cells, np.int32)
        mir = np.zeros(cap, np.int32)
        pos = np.zeros(2 * nf, np.int32)
        nnz = self.lib.tfs_ref_topology(n_cells, n_bnd, nf, faces_c, cap, row, col, dia, mir, pos)
        return dict(row=row, col=col[:nnz], dia=dia, mir=mir[:nnz], pos=pos)

    def pressure_matrix(self, n_cells, faces_c, dt, scoef, n_bnd=None):
        """Pressure_Matrix_Fractional on the topology of faces_c: A % val (0-based array)."""
        faces_c = _c32(faces_c)
        nf = faces_c.size // 2
        if n_bnd is None:
            n_bnd = int(max(0, -faces_c[1::2].min())) if nf else 0
        cap = n_cells + 2 * nf
        val = np.zeros(cap, np.float64)
        nnz = self.lib.tfs_ref_pressure_matrix(n_cells, n_bnd, nf, faces_c, float(dt),
                                               _c64(scoef), cap, val)
        return val[:nnz]

    # what tfs_ref_mesh_get returns, by code: (name, integer array?)
    MESH_ARRAYS = {1: ("faces_c", True), 2: ("row", True), 3: ("col", True), 4: ("dia", True),
                   5: ("xc", False), 6: ("vol", False), 7: ("s", False), 8: ("shift", False),
                   9: ("f", False), 10: ("xf", False), 11: ("d", False), 12: ("scoef", False),
                   13: ("val", False), 14: ("bcmark", True), 15: ("copy_c", True)}

    def mesh(self, stem, dt=1.0):
        """The reference's Load_Cns + Load_Geo on <stem>.cns / <stem>.geo,
        Compute_Geometry (every boundary a wall), Matrix_Mod_Topology and
        Pressure_Matrix_Fractional with time step dt.  Returns dict of the
        arrays listed in MESH_ARRAYS (row/col/dia 1-based, as the reference
        holds them) plus the sizes n_cells, n_bnd_cells, n_faces, nnz.  The
        reference reads from the working directory, so the call runs there."""
        import os
        stem = pathlib.Path(stem).resolve()
        sizes = np.zeros(4, np.int32)
        here = os.getcwd()
        os.chdir(stem.parent)
        try:
            name = stem.name.encode()
            rc = self.lib.tfs_ref_mesh_load(name, len(name), float(dt), sizes)
        finally:
            os.chdir(here)
        if rc != 0:
            raise RuntimeError(f"tfs_ref_mesh_load({stem}) returned {rc}")
        nc, nb, nf, nnz = (int(v) for v in sizes)
        cap = max(nc + nb, nf, nnz, nc + 1) * 3
        out = dict(n_cells=nc, n_bnd_cells=nb, n_faces=nf, nnz=nnz)
        ibuf = np.zeros(cap, np.int32)
        rbuf = np.zeros(cap, np.float64)
        for code, (key, is_int) in self.MESH_ARRAYS.items():
            k = self.lib.tfs_ref_mesh_get(code, cap, ibuf.ctypes.data, rbuf.ctypes.data)
            if k < 0:
                raise RuntimeError(f"tfs_ref_mesh_get({code}) returned {k}")
            out[key] = (ibuf if is_int else rbuf)[:k].copy()
        return out

    def mesh_grad(self, mode, n_cells, n_bnd, phi=None, bc=None):
        """On the grid of the last mesh() call: mode 0 -> G (n_cells, 6) from
        Compute_Gradient_Matrix; mode 1 -> (px, py, pz) from GraPhi; mode 2 ->
        (phi, px, py, pz) from GradP.  bc: kind per boundary cell (0 wall,
        1 symmetry, 2 pressure, 3 buffer), default walls."""
        bcb = np.zeros(max(n_bnd, 1), np.int8) if bc is None else \
            np.ascontiguousarray(bc, np.int8)
        f = np.zeros(n_cells + n_bnd) if phi is None else _c64(phi).copy()
        px, py, pz = (np.zeros(n_cells) for _ in range(3))
        g = np.zeros((n_cells, 6))
        rc = self.lib.tfs_ref_mesh_grad(mode, bcb.ctypes.data, f.ctypes.data, px.ctypes.data,
                                        py.ctypes.data, pz.ctypes.data, g.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"tfs_ref_mesh_grad({mode}) returned {rc}")
        return g if mode == 0 else (px, py, pz) if mode == 1 else (f, px, py, pz)

    def mesh_correct(self, n_cells, u, v, w, px, py, pz, pp, sav, flux, den, visc, bc=None):
        """Correct_Velocity on the grid and pressure matrix of the last mesh()
        call (material 1 everywhere): dict(u, v, w, flux, b, errmax, cfl, pe),
        cfl and pe to the four digits the reference prints."""
        n_bnd = len(u) - n_cells
        bcb = np.zeros(max(n_bnd, 1), np.int8) if bc is None else np.ascontiguousarray(bc, np.int8)
        u, v, w, flux = (_c64(a).copy() for a in (u, v, w, flux))
        keep = [_c64(a) for a in (px, py, pz, pp, sav)]
        b = np.zeros(n_cells)
        res = np.zeros(3)
        rc = self.lib.tfs_ref_mesh_correct(bcb.ctypes.data, u.ctypes.data, v.ctypes.data,
                                           w.ctypes.data, *(a.ctypes.data for a in keep),
                                           flux.ctypes.data, float(den), float(visc),
                                           b.ctypes.data, res.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"tfs_ref_mesh_correct returned {rc}")
        return dict(u=u, v=v, w=w, flux=flux, b=b, errmax=res[0], cfl=res[1], pe=res[2])

    def mesh_press_simple(self, n_cells, n_faces, nnz, u, v, w, p, px, py, pz, sav, den, bc,
                          prec, stol, urf):
        """Compute_Pressure_Simple on the grid and matrix of the last mesh()
        call with boundary-cell kinds bc (0 wall, 1 symmetry, 2 pressure,
        4 inflow, 5 outflow, 6 convective): dict(flux, val, pp, p)."""
        keep = [_c64(a) for a in (u, v, w, p, px, py, pz, sav)]
        bck = np.ascontiguousarray(bc, np.int32)
        flux, val = np.zeros(n_faces), np.zeros(nnz)
        pp, pn = np.zeros(n_cells), np.zeros(n_cells)
        rc = self.lib.tfs_ref_mesh_press_simple(*(a.ctypes.data for a in keep), float(den),
                                                bck.ctypes.data, int(prec), float(stol),
                                                float(urf), flux.ctypes.data, val.ctypes.data,
                                                pp.ctypes.data, pn.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"tfs_ref_mesh_press_simple returned {rc}")
        return dict(flux=flux, val=val, pp=pp, p=pn)

    def mesh_press_frac(self, n_cells, n_faces, nnz, u, v, w, p, px, py, pz, sav, den, prec,
                        stol, urf):
        """Compute_Pressure_Fractional on the grid and matrix of the last mesh()
        call: dict(flux, val, pp, p) -- face fluxes, A % val, the pressure
        correction PP % n and the updated pressure P % n (cells)."""
        keep = [_c64(a) for a in (u, v, w, p, px, py, pz, sav)]
        flux, val = np.zeros(n_faces), np.zeros(nnz)
        pp, pn = np.zeros(n_cells), np.zeros(n_cells)
        rc = self.lib.tfs_ref_mesh_press_frac(*(a.ctypes.data for a in keep), float(den),
                                              int(prec), float(stol), float(urf), flux.ctypes.data,
                                              val.ctypes.data, pp.ctypes.data, pn.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"tfs_ref_mesh_press_frac returned {rc}")
        return dict(flux=flux, val=val, pp=pp, p=pn)
