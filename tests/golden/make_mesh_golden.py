"""Writes tests/golden/mesh/: real T-FlowS meshes and what the REFERENCE makes
of them.

Run in a container that has the reference tree.  ``make -C oracle ref mesh``
compiles the reference's mesh generator and its solver path from their own
sources and runs the generator on its Tests/ domain files (cavity, and the
non-orthogonal backstep).  This script then keeps, per mesh:

  * <name>.cns.gz / <name>.geo.gz  -- the generator's output files, as data;
  * in mesh_reference.json: the sizes and the SHA-256 of every array the
    reference's Load_Cns / Load_Geo read, of Compute_Geometry's distances and
    Scoef, and of Matrix_Mod_Topology's row/col/dia and
    Pressure_Matrix_Fractional's val at time step DT (little-endian int32 /
    float64 bytes, 1-based indices as the reference holds them);
  * in mesh_reference.json, "gradients": the SHA-256 of the reference's
    Compute_Gradient_Matrix G, of GraPhi's three components and of GradP's
    (phi, px, py, pz) for the seeded field and boundary kinds of
    tests/mesh_fixtures.grad_inputs (every boundary a wall for "walls", mixed kinds for
    "mixed");
  * in mesh_reference.json, "correct_velocity": the SHA-256 of the
    reference's Correct_Velocity outputs (u, v, w, flux, b) for the seeded
    inputs of tests/mesh_fixtures.cv_inputs on the pressure matrix above,
    its errmax (exact, as a float hex string) and the CFL / Peclet maxima
    as it prints them (four digits);
  * in mesh_reference.json, "pressure_fractional": the SHA-256 of the
    reference's Compute_Pressure_Fractional outputs for the same inputs (pp
    as the old pressure): face fluxes, A % val, PP % n and P % n;
  * in mesh_reference.json, "pressure_simple": the same for the reference's
    Compute_Pressure_Simple with the boundary kinds of
    tests/mesh_fixtures.simple_kinds (walls, symmetry, pressure, inflow,
    outflow and convective boundaries; Bicg prec 2 to 1e-9, cap 40);
  * in <name>_cg_p2.npz: the reference's Cg (prec 2) on that pressure matrix
    with cell 1 pinned (a_11 doubled) and the seeded right-hand side of
    tflows_amd.mesh.mesh_case, to tol 1e-10 ||b||: x, niter, ini_res, fin_res.

    python tests/golden/make_mesh_golden.py
"""
import gzip
import hashlib
import json
import pathlib
import shutil
import subprocess
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))
from conftest import load_package  # noqa: E402

load_package()
from tflows_amd import mesh  # noqa: E402
from oracle.oracle import Reference  # noqa: E402
from mesh_fixtures import CV_DEN, CV_VISC, SEED, cv_inputs, grad_inputs, simple_kinds  # noqa: E402,E501

OUT = HERE / "mesh"
GEN = ROOT / "oracle" / "_ref" / "gen"
MESHES = ["cavity", "backstep_nonorthogonal"]
DT = 0.013


def sha(a):
    a = np.ascontiguousarray(a)
    a = a.astype(a.dtype.newbyteorder("<"), copy=False)
    return hashlib.sha256(a.tobytes()).hexdigest()


def main():
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref", "mesh"], check=True)
    OUT.mkdir(exist_ok=True)
    ref = Reference()
    meta = {"dt": DT, "meshes": {}}
    for name in MESHES:
        stem = GEN / name / name
        for ext in (".cns", ".geo"):
            with open(str(stem) + ext, "rb") as src, \
                    gzip.GzipFile(OUT / (name + ext + ".gz"), "wb", 9, mtime=0) as dst:
                shutil.copyfileobj(src, dst)
        r = ref.mesh(stem, dt=DT)
        entry = {k: r[k] for k in ("n_cells", "n_bnd_cells", "n_faces", "nnz")}
        entry["sha256"] = {key: sha(r[key]) for key, _ in Reference.MESH_ARRAYS.values()}
        nc, nb = r["n_cells"], r["n_bnd_cells"]
        grads = {"G": sha(ref.mesh_grad(0, nc, nb))}
        for tag, mixed in (("walls", False), ("mixed", True)):
            phi, bc = grad_inputs(nc, nb, mixed)
            grads[f"graphi_{tag}"] = [sha(a) for a in ref.mesh_grad(1, nc, nb, phi, bc)]
            grads[f"gradp_{tag}"] = [sha(a) for a in ref.mesh_grad(2, nc, nb, phi, bc)]
        entry["gradients"] = grads
        # Correct_Velocity on the pressure matrix just built (material 1)
        cv = cv_inputs(nc, nb, r["n_faces"])
        out = ref.mesh_correct(nc, cv["u"], cv["v"], cv["w"], cv["px"], cv["py"], cv["pz"],
                               cv["pp"], cv["sav"], cv["flux"], CV_DEN, CV_VISC, cv["bc"])
        entry["correct_velocity"] = {k: sha(out[k]) for k in ("u", "v", "w", "flux", "b")}
        entry["correct_velocity"].update(errmax=float(out["errmax"]).hex(),
                                         cfl=float(out["cfl"]), pe=float(out["pe"]))
        # Compute_Pressure_Fractional (fluid of density CV_DEN, walls, Cg prec 2
        # to 1e-9 with the routine's cap 200, pressure under-relaxation 0.8)
        pf = ref.mesh_press_frac(nc, r["n_faces"], r["nnz"], cv["u"], cv["v"], cv["w"],
                                 cv["pp"], cv["px"], cv["py"], cv["pz"], cv["sav"], CV_DEN, 2,
                                 1e-9, 0.8)
        entry["pressure_fractional"] = {k: sha(pf[k]) for k in ("flux", "val", "pp", "p")}
        # Compute_Pressure_Simple (the same fields, mixed boundary kinds)
        ps = ref.mesh_press_simple(nc, r["n_faces"], r["nnz"], cv["u"], cv["v"], cv["w"],
                                   cv["pp"], cv["px"], cv["py"], cv["pz"], cv["sav"], CV_DEN,
                                   simple_kinds(nb), 2, 1e-9, 0.8)
        entry["pressure_simple"] = {k: sha(ps[k]) for k in ("flux", "val", "pp", "p")}
        meta["meshes"][name] = entry

        # the pinned pressure system, solved by the reference
        row, col, dia = r["row"] - 1, r["col"] - 1, r["dia"] - 1
        val = r["val"].copy()
        val[dia[0]] *= 2.0
        c = mesh.mesh_case(stem, dt=DT, seed=SEED)
        assert np.array_equal(c.val, val) and np.array_equal(c.col, col)
        tol = 1e-10 * float(np.linalg.norm(c.b))
        ref.set_matrix(row, col, dia, val)
        out = ref.solve("cg", c.b, prec=2, niter=20000, tol=tol)
        np.savez_compressed(OUT / f"{name}_cg_p2.npz", x=out["x"], niter=out["niter"],
                            ini_res=out["ini_res"], fin_res=out["fin_res"], tol=tol,
                            b_sha256=sha(c.b), seed=SEED, dt=DT)
        print(name, entry["n_cells"], "cells; Cg prec 2:", out["niter"], "iterations")
    (OUT / "mesh_reference.json").write_text(json.dumps(meta, indent=1) + "\n")


if __name__ == "__main__":
    main()
