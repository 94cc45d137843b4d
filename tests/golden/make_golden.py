"""Writes tests/golden/*.npz: outputs of the REFERENCE solver path.

Run in a container that has the reference tree: ``make -C oracle ref`` builds
oracle/_ref/libtfs_ref.so from the reference's own Fortran sources (its FAST
build flags), and this script calls that library on seeded synthetic systems.
Inputs are not stored: they are regenerated from the recorded case parameters
(tflows_amd.mesh.make_case), and the SHA-256 of row/col/dia/val/b is stored so
that a changed generator is detected.  Outputs stored per solve: x, r (the
final residual the reference leaves in b), niter, ini_res, fin_res.

    python tests/golden/make_golden.py [key prefix]
"""
import hashlib
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))
from conftest import load_package  # noqa: E402

load_package()
from tflows_amd import cases, mesh  # noqa: E402
from oracle.oracle import Reference  # noqa: E402


def digest(c):
    h = hashlib.sha256()
    for a in (c.row, c.col, c.dia, c.val, c.b):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


SMALL = {
    "poisson12": dict(kind=mesh.POISSON, nx=12, beta=1.5),
    "momentum10": dict(kind=mesh.MOMENTUM, nx=10, beta=2.0, jitter=0.1, permute=True,
                       pin_first=False, rhs_zero_mean=False),
    "poisson_aniso": dict(kind=mesh.POISSON, nx=9, ny=7, nz=5, beta=(0.0, 2.0, 3.0),
                          permute=True),
}


def main():
    R = Reference()
    out = {}
    # small systems, every method x preconditioner, relative 1e-10
    for name, kw in SMALL.items():
        c = mesh.make_case(**kw)
        R.set_matrix(c.row, c.col, c.dia, c.val)
        tol = 1e-10 * float(np.linalg.norm(c.b))
        for meth in ("cg", "bicg", "cgs"):
            for prec in (0, 1, 2):
                r = R.solve(meth, c.b, prec=prec, niter=2000, tol=tol)
                key = f"{name}_{meth}_p{prec}"
                out[key] = dict(case=kw, sha256=digest(c), method=meth, prec=prec, tol=tol,
                                cap=2000, niter=r["niter"], ini_res=r["ini_res"],
                                fin_res=r["fin_res"], x=r["x"], r=r["r"])
        # preconditioner pieces
        for prec in (1, 2):
            z, d = R.prec(prec, c.b)
            out[f"{name}_prec{prec}"] = dict(case=kw, sha256=digest(c), prec=prec, z=z, d=d)
        rr = R.residual(c.b, c.b)
        out[f"{name}_residual"] = dict(case=kw, sha256=digest(c), r=rr)
    # the 32^3 Poisson workload: CG, DIC, relative 1e-8, cap 1e4
    kw1 = cases.workload_kwargs(1)
    c = mesh.make_case(**kw1)
    R.set_matrix(c.row, c.col, c.dia, c.val)
    tol = 1e-8 * float(np.linalg.norm(c.b))
    r = R.solve("cg", c.b, prec=2, niter=10000, tol=tol)
    out["poisson32_cg_p2"] = dict(case=kw1, sha256=digest(c), method="cg", prec=2, tol=tol,
                                cap=10000, niter=r["niter"], ini_res=r["ini_res"],
                                fin_res=r["fin_res"], x=r["x"], r=r["r"])
    # topology of a permuted jittered mesh (Matrix_Mod_Topology)
    kwt = dict(kind=mesh.MOMENTUM, nx=6, ny=5, nz=4, beta=2.0, jitter=0.1, permute=True)
    c = mesh.make_case(with_faces=True, **kwt)
    t = R.topology(c.n, c.faces)
    out["topology_6x5x4"] = dict(case=kwt, n=c.n, **t)
    # pressure matrix of the same mesh (Pressure_Matrix_Fractional), seeded
    # face coefficients Scoef and dt as inputs
    scoef = np.random.default_rng(20261202).uniform(0.1, 2.0, c.faces.size // 2)
    dt = 0.0137
    out["pressure_matrix_6x5x4"] = dict(case=kwt, n=c.n, dt=dt, scoef=scoef,
                                        val=R.pressure_matrix(c.n, c.faces, dt, scoef))

    only = sys.argv[1] if len(sys.argv) > 1 else ""
    for key, rec in out.items():
        if not key.startswith(only):
            continue
        arrays = {k: v for k, v in rec.items() if isinstance(v, np.ndarray)}
        meta = {k: v for k, v in rec.items() if not isinstance(v, np.ndarray)}
        np.savez_compressed(HERE / f"{key}.npz", meta=np.array(json.dumps(meta)), **arrays)
    print(f"wrote {len(out)} fixtures to {HERE}")


if __name__ == "__main__":
    main()
