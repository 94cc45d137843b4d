"""Synthetic hex-mesh cases and the CRS topology builder (libtfs_mesh.so).

``make_case`` returns the reference's matrix layout (row/col/dia/val, 0-based
here; ``base=1`` gives the Fortran numbering) and the right-hand side.  The
matrices follow the reference's assembly formulas:

* Poisson (fractional-step pressure): off-diagonal ``-Scoef`` per interior face,
  diagonal the sum (Sources/Process/Pressure_Matrix_Fractional.f90:26-48,
  Sources/Process/Compute_Geometry.f90:22-49); Neumann walls; optionally cell 1
  pinned by doubling ``a_11``.
* Momentum: fully implicit diffusion with upwind convection
  ``A12 = A0 - min(F,0)``, ``A21 = A0 + max(F,0)`` and an inertia diagonal
  ``V/dt`` (Sources/Process/NewUVW.f90:349-384,500-506), with face fluxes from a
  seeded divergence-free field.

``mesh_case`` builds the same Poisson matrix on a real T-FlowS mesh read from
its ``.cns`` / ``.geo`` files (``read_cns``, ``read_geo``, ``face_coef``,
``pressure_values``).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native

POISSON = 0
MOMENTUM = 1


class _Params(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("beta", ctypes.c_double * 3),
        ("jitter", ctypes.c_double),
        ("permute", ctypes.c_int),
        ("pin_first", ctypes.c_int),
        ("rhs_zero_mean", ctypes.c_int),
        ("nu", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("n_modes", ctypes.c_int),
        ("seed", ctypes.c_uint64),
    ]


_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _lib():
    lib = _native.mesh_lib()
    if not getattr(lib, "_tfs_typed", False):
        lib.tfs_rng_u01.restype = ctypes.c_double
        lib.tfs_rng_u01.argtypes = [ctypes.c_uint64] * 3
        lib.tfs_topology_nnz.restype = ctypes.c_int64
        lib.tfs_topology_nnz.argtypes = [ctypes.c_int, ctypes.c_int64, _I32P]
        lib.tfs_topology_build.restype = ctypes.c_int
        lib.tfs_topology_build.argtypes = [ctypes.c_int, ctypes.c_int64, _I32P, ctypes.c_int,
                                           _I32P, _I32P, _I32P, ctypes.c_void_p]
        lib.tfs_case_create.restype = ctypes.c_int
        lib.tfs_case_create.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(ctypes.c_void_p)]
        lib.tfs_case_destroy.argtypes = [ctypes.c_void_p]
        lib.tfs_case_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_int64)]
        lib.tfs_case_crs.restype = ctypes.c_int
        lib.tfs_case_crs.argtypes = [ctypes.c_void_p, ctypes.c_int, _I32P, _I32P, _I32P,
                                     _F64P, _F64P]
        lib.tfs_case_faces.restype = ctypes.c_int
        lib.tfs_case_faces.argtypes = [ctypes.c_void_p, _I32P]
        lib.tfs_case_face_coefs.restype = ctypes.c_int
        lib.tfs_case_face_coefs.argtypes = [ctypes.c_void_p, _F64P, _F64P, _F64P]
        lib.tfs_case_dt.restype = ctypes.c_double
        lib.tfs_case_dt.argtypes = [ctypes.c_void_p]
        lib.tfs_case_max_aspect.restype = ctypes.c_double
        lib.tfs_case_max_aspect.argtypes = [ctypes.c_void_p]
        lib.tfs_layout_colour.restype = ctypes.c_int
        lib.tfs_layout_colour.argtypes = [ctypes.c_int, _I32P, _I32P, _I32P, _I32P]
        lib.tfs_layout_block_colour.restype = ctypes.c_int
        lib.tfs_layout_block_colour.argtypes = [ctypes.c_int, _I32P, _I32P, ctypes.c_int, _I32P,
                                                _I32P]
        lib.tfs_layout_brick_colour.restype = ctypes.c_int
        lib.tfs_layout_brick_colour.argtypes = [ctypes.c_int, _I32P, _I32P, _I32P, ctypes.c_int,
                                                ctypes.c_int, _I32P, _I32P, _I32P]
        lib.tfs_layout_block_colour_bfs.restype = ctypes.c_int
        lib.tfs_layout_block_colour_bfs.argtypes = [ctypes.c_int, _I32P, _I32P, ctypes.c_int,
                                                    _I32P, _I32P, _I32P]
        lib.tfs_mesh_read_cns.restype = ctypes.c_int
        lib.tfs_mesh_read_cns.argtypes = [ctypes.c_char_p] + [ctypes.POINTER(ctypes.c_int)] * 3 + \
            [ctypes.c_void_p, ctypes.c_void_p]
        lib.tfs_mesh_read_cns_marks.restype = ctypes.c_int
        lib.tfs_mesh_read_cns_marks.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_void_p]
        lib.tfs_mesh_read_geo.restype = ctypes.c_int
        lib.tfs_mesh_read_geo.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int] + [ctypes.c_void_p] * 6
        lib.tfs_mesh_face_coef.restype = ctypes.c_int
        lib.tfs_mesh_face_coef.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _I32P, _F64P,
                                           _F64P, _F64P, ctypes.c_void_p, _F64P]
        lib.tfs_mesh_pressure_values.restype = ctypes.c_int
        lib.tfs_mesh_pressure_values.argtypes = [ctypes.c_int, ctypes.c_int64, _I32P, _I32P,
                                                 _I32P, ctypes.c_int, ctypes.c_int64, _F64P,
                                                 ctypes.c_double, ctypes.c_void_p, _F64P]
        lib.tfs_layout_summary.restype = ctypes.c_int
        lib.tfs_layout_summary.argtypes = [ctypes.c_int, _I32P, _I32P, _I32P, _I32P, _I32P,
                                           _I32P, ctypes.POINTER(ctypes.c_int)]
        lib._tfs_typed = True
    return lib


def rng_u01(seed, tag, i):
    """i-th uniform draw in [0,1) of the SplitMix64 stream keyed by (seed, tag)."""
    return _lib().tfs_rng_u01(seed, tag, i)


@dataclass
class Case:
    """One linear system in the reference CRS layout (index base ``base``)."""
    n: int
    row: np.ndarray
    col: np.ndarray
    dia: np.ndarray
    val: np.ndarray
    b: np.ndarray
    base: int = 0
    dt: float = 0.0
    max_aspect: float = 1.0
    faces: np.ndarray = None
    face_coefs: tuple = None   # (a12, a21, dextra) when built with_faces
    params: dict = field(default_factory=dict)

    @property
    def nnz(self):
        return int(self.row[-1] - self.base)


def make_case(kind=POISSON, nx=8, ny=None, nz=None, beta=(0.0, 0.0, 0.0), jitter=0.0,
              permute=False, pin_first=True, rhs_zero_mean=True, nu=1e-3, cfl=1.0,
              n_modes=8, seed=20261202, base=0, with_faces=False):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    if np.isscalar(beta):
        beta = (float(beta),) * 3
    p = _Params(kind=kind, nx=nx, ny=ny, nz=nz, beta=(ctypes.c_double * 3)(*beta),
                jitter=jitter, permute=int(permute), pin_first=int(pin_first),
                rhs_zero_mean=int(rhs_zero_mean), nu=nu, cfl=cfl, n_modes=n_modes,
                seed=seed)
    lib = _lib()
    h = ctypes.c_void_p()
    rc = lib.tfs_case_create(ctypes.byref(p), ctypes.byref(h))
    if rc != 0:
        raise ValueError(f"tfs_case_create failed ({rc})")
    try:
        n = ctypes.c_int()
        nf = ctypes.c_int64()
        nnz = ctypes.c_int64()
        lib.tfs_case_sizes(h, ctypes.byref(n), ctypes.byref(nf), ctypes.byref(nnz))
        n, nf, nnz = n.value, nf.value, nnz.value
        row = np.empty(n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        dia = np.empty(n, np.int32)
        val = np.empty(nnz, np.float64)
        b = np.empty(n, np.float64)
        rc = lib.tfs_case_crs(h, base, row, col, dia, val, b)
        if rc != 0:
            raise ValueError(f"tfs_case_crs failed ({rc})")
        faces = coefs = None
        if with_faces:
            faces = np.empty(2 * nf, np.int32)
            lib.tfs_case_faces(h, faces)
            coefs = (np.empty(nf), np.empty(nf), np.empty(n))
            lib.tfs_case_face_coefs(h, *coefs)
        params = dict(kind=kind, nx=nx, ny=ny, nz=nz, beta=tuple(beta), jitter=jitter,
                      permute=bool(permute), pin_first=bool(pin_first),
                      rhs_zero_mean=bool(rhs_zero_mean), nu=nu, cfl=cfl, n_modes=n_modes,
                      seed=seed)
        return Case(n=n, row=row, col=col, dia=dia, val=val, b=b, base=base,
                    dt=lib.tfs_case_dt(h), max_aspect=lib.tfs_case_max_aspect(h),
                    faces=faces, face_coefs=coefs, params=params)
    finally:
        lib.tfs_case_destroy(h)


def topology(n_cells, faces_c, base=0, with_pos=False):
    """CRS pattern from a face list (faces_c: flat (2, n_faces) column-major,
    1-based cells, c2 <= 0 boundary).  Matrix_Mod_Topology counterpart."""
    lib = _lib()
    faces_c = np.ascontiguousarray(faces_c, dtype=np.int32)
    nf = faces_c.size // 2
    nnz = lib.tfs_topology_nnz(n_cells, nf, faces_c)
    row = np.empty(n_cells + 1, np.int32)
    col = np.empty(nnz, np.int32)
    dia = np.empty(n_cells, np.int32)
    pos = np.empty(2 * nf, np.int32) if with_pos else None
    rc = lib.tfs_topology_build(n_cells, nf, faces_c, base, row, col, dia,
                                pos.ctypes.data if with_pos else None)
    if rc != 0:
        raise ValueError(f"tfs_topology_build failed ({rc})")
    return (row, col, dia, pos) if with_pos else (row, col, dia)


def layout_colour(row, col, block=1):
    """Greedy (block) colouring the FP32 inner solves use: (ncolours, perm, colour)."""
    row = np.ascontiguousarray(row, np.int32)
    n = len(row) - 1
    perm = np.empty(n, np.int32)
    colour = np.empty(n, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    if block == 1:
        k = _lib().tfs_layout_colour(n, row, col, perm, colour)
    else:
        k = _lib().tfs_layout_block_colour(n, row, col, int(block), perm, colour)
    if k < 0:
        raise ValueError("tfs_layout_colour failed")
    return k, perm, colour


def layout_block_colour_bfs(row, col, block):
    """Block multi-colour order of the FP32 inner solves (blocks of at most
    `block` rows grown breadth-first): (ncolours, perm, colour, block_of)."""
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    n = len(row) - 1
    perm, colour, block_of = (np.empty(n, np.int32) for _ in range(3))
    k = _lib().tfs_layout_block_colour_bfs(n, row, col, int(block), perm, colour, block_of)
    if k < 0:
        raise ValueError("tfs_layout_block_colour_bfs failed")
    return k, perm, colour, block_of


def layout_brick_colour(row, col, dia, edge=4, cap=64):
    """Brick order of the FP32 inner solves (TFS_INNER_BRICKS): blocks from
    the monotone potentials of the caller's lower DAG, coloured greedily:
    (ncolours, perm, colour, block_of)."""
    row = np.ascontiguousarray(row, np.int32)
    n = len(row) - 1
    perm, colour, block_of = (np.empty(n, np.int32) for _ in range(3))
    k = _lib().tfs_layout_brick_colour(n, row, np.ascontiguousarray(col, np.int32),
                                       np.ascontiguousarray(dia, np.int32), int(edge), int(cap),
                                       perm, colour, block_of)
    if k < 0:
        raise ValueError("tfs_layout_brick_colour failed")
    return k, perm, colour, block_of


def layout_summary(row, col, dia):
    """The solver's layouts of a 0-based CRS matrix: dict(s_orig, level_lower,
    level_upper, bandwidth)."""
    row = np.ascontiguousarray(row, np.int32)
    n = len(row) - 1
    s_orig = np.empty(n, np.int32)
    ll = np.empty(n, np.int32)
    lu = np.empty(n, np.int32)
    bw = ctypes.c_int()
    rc = _lib().tfs_layout_summary(n, row, np.ascontiguousarray(col, np.int32),
                                   np.ascontiguousarray(dia, np.int32), s_orig, ll, lu,
                                   ctypes.byref(bw))
    if rc != 0:
        raise ValueError("tfs_layout_summary failed")
    return dict(s_orig=s_orig, level_lower=ll, level_upper=lu, bandwidth=bw.value)


# ---- T-FlowS mesh files ------------------------------------------------------
MESH_ERRORS = {-1: "invalid argument", -2: "cannot open the file", -3: "bad record"}
TAG_MESH_RHS = 0x4D455348   # stream of the right-hand side of mesh_case


def _mesh_check(rc, what):
    if rc != 0:
        raise ValueError(f"{what}: {MESH_ERRORS.get(rc, rc)}")


def read_cns(path):
    """Connectivity of a T-FlowS .cns file (Sources/Process/Load_Cns.f90:25-69):
    dict(n_cells, n_bnd_cells, n_faces, faces_c, material).  faces_c is the
    reference's faces_c(2, n_faces) flattened column-major (c1, c2 per face),
    1-based, c2 = -1..-n_bnd_cells for boundary faces; material lists the
    cells, then the boundary cells; bcmark / copy_c (Load_Cns.f90:62-69) the
    boundary-condition number and the copy cell of every boundary cell."""
    lib = _lib()
    nc, nb, nf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    p = str(path).encode()
    _mesh_check(lib.tfs_mesh_read_cns(p, ctypes.byref(nc), ctypes.byref(nb), ctypes.byref(nf),
                                      None, None), f"read_cns({path})")
    faces = np.empty(2 * nf.value, np.int32)
    material = np.empty(nc.value + nb.value, np.int32)
    _mesh_check(lib.tfs_mesh_read_cns(p, ctypes.byref(nc), ctypes.byref(nb), ctypes.byref(nf),
                                      faces.ctypes.data, material.ctypes.data),
                f"read_cns({path})")
    bcmark = np.empty(max(nb.value, 1), np.int32)
    copy_c = np.empty(max(nb.value, 1), np.int32)
    _mesh_check(lib.tfs_mesh_read_cns_marks(p, nb.value, bcmark.ctypes.data, copy_c.ctypes.data),
                f"read_cns({path})")
    return dict(n_cells=nc.value, n_bnd_cells=nb.value, n_faces=nf.value, faces_c=faces,
                material=material, bcmark=bcmark[:nb.value], copy_c=copy_c[:nb.value])


def read_geo(path, cns):
    """Geometry of a .geo file for its .cns (Sources/Shared/Load_Geo.f90:23-50):
    dict(xc (3, n_cells + n_bnd_cells) -- cells, then boundary cells --, vol,
    s (3, n_faces) surface vectors, shift (3, n_faces) periodic shifts, f
    weights, xf (3, n_faces) face centres)."""
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    out = dict(xc=np.empty((3, nc + nb)), vol=np.empty(nc), s=np.empty((3, nf)),
               shift=np.empty((3, nf)), f=np.empty(nf), xf=np.empty((3, nf)))
    _mesh_check(_lib().tfs_mesh_read_geo(str(path).encode(), nc, nb, nf,
                                         *(out[k].ctypes.data for k in
                                           ("xc", "vol", "s", "shift", "f", "xf"))),
                f"read_geo({path})")
    return out


def face_coef(cns, geo):
    """(scoef, d): Scoef = |S|^2 / (d . S) per face and the centre-to-centre
    vectors d (3, n_faces), Sources/Process/Compute_Geometry.f90:22-49."""
    nf = cns["n_faces"]
    scoef = np.empty(nf)
    d = np.empty((3, nf))
    _mesh_check(_lib().tfs_mesh_face_coef(cns["n_cells"], cns["n_bnd_cells"], nf, cns["faces_c"],
                                          geo["xc"].reshape(-1), geo["s"].reshape(-1),
                                          geo["shift"].reshape(-1), d.ctypes.data, scoef),
                "face_coef")
    return scoef, d


def pressure_values(n_cells, faces_c, pos, dia, nnz, scoef, dt=1.0, bnd_on=None, base=0):
    """Values of the fractional-step pressure matrix on a topology from
    ``topology(..., with_pos=True)``: Pressure_Matrix_Fractional.f90:19-48."""
    faces_c = np.ascontiguousarray(faces_c, np.int32)
    val = np.empty(int(nnz))
    bnd = None if bnd_on is None else np.ascontiguousarray(bnd_on, np.uint8)
    _mesh_check(_lib().tfs_mesh_pressure_values(
        int(n_cells), faces_c.size // 2, faces_c, np.ascontiguousarray(pos, np.int32),
        np.ascontiguousarray(dia, np.int32), int(base), int(nnz),
        np.ascontiguousarray(scoef, np.float64), float(dt),
        None if bnd is None else bnd.ctypes.data, val), "pressure_values")
    return val


def mesh_case(stem, dt=1.0, pin_first=True, seed=20261202, base=0):
    """The fractional-step pressure system on the T-FlowS mesh <stem>.cns /
    <stem>.geo: topology as Matrix_Mod_Topology builds it, values as
    Pressure_Matrix_Fractional forms them (every boundary a wall), cell 1
    pinned by doubling a_11 when pin_first, and a seeded zero-mean uniform
    right-hand side.  face_coefs holds (a12 = a21 = dt Scoef on interior
    faces and 0 on boundary faces, no cell term): assembled on the device they
    give the values before pinning."""
    stem = str(stem)
    cns = read_cns(stem + ".cns")
    geo = read_geo(stem + ".geo", cns)
    scoef, _ = face_coef(cns, geo)
    n, faces = cns["n_cells"], cns["faces_c"]
    row, col, dia, pos = topology(n, faces, base=0, with_pos=True)
    val = pressure_values(n, faces, pos, dia, len(col), scoef, dt)
    if pin_first:
        val[dia[0]] *= 2.0
    lib = _lib()
    b = np.array([2 * lib.tfs_rng_u01(seed, TAG_MESH_RHS, i) - 1 for i in range(n)])
    b -= b.mean()
    if base:
        row, col, dia = row + base, col + base, dia + base
    a = dt * scoef
    a[faces[1::2] <= 0] = 0.0      # walls add nothing to the diagonal
    return Case(n=n, row=row, col=col, dia=dia, val=val, b=b, base=base, dt=dt, faces=faces,
                face_coefs=(a, a.copy(), np.zeros(n)),
                params=dict(mesh=stem, dt=dt, pin_first=pin_first, seed=seed,
                            n_bnd_cells=cns["n_bnd_cells"], n_faces=cns["n_faces"]))
