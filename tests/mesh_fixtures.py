"""Shared by the mesh tests and tests/golden/make_mesh_golden.py: the
committed reference meshes and the seeded inputs of the gradient checks."""
import gzip
import json
import pathlib
import shutil

import numpy as np

from tflows_amd import mesh

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "mesh"
SEED = 20261202


def meta():
    return json.loads((GOLDEN / "mesh_reference.json").read_text())


def unpack(dest):
    """Decompresses the fixture meshes into dest; returns their names."""
    names = list(meta()["meshes"])
    for name in names:
        for ext in (".cns", ".geo"):
            with gzip.open(GOLDEN / (name + ext + ".gz"), "rb") as src, \
                    open(pathlib.Path(dest) / (name + ext), "wb") as dst:
                shutil.copyfileobj(src, dst)
    return names


def grad_inputs(n_cells, n_bnd, mixed):
    """A seeded field on cells and boundary cells (uniform in [-1, 1)) and
    boundary kinds (0 wall, 1 symmetry, 2 pressure) from the SplitMix64
    stream of tflows_amd.mesh, so the inputs are exact on any machine."""
    lib = mesh._lib()
    phi = np.array([2.0 * lib.tfs_rng_u01(SEED, 0x47524144, i) - 1.0
                    for i in range(n_cells + n_bnd)])
    if not mixed:
        return phi, np.zeros(n_bnd, np.int8)
    bc = np.array([int(3.0 * lib.tfs_rng_u01(SEED, 0x4B494E44, i)) for i in range(n_bnd)],
                  np.int8)
    return phi, bc


CV_DEN, CV_VISC = 1.3, 1e-3


def cv_inputs(n_cells, n_bnd, n_faces):
    """Seeded inputs of the Correct_Velocity checks: u, v, w, pp on cells and
    boundary cells and px, py, pz on cells uniform in [-1, 1); sav uniform in
    [0.5, 1.5); fluxes uniform in [-1, 1); boundary kinds mixed (0 wall,
    1 symmetry, 2 pressure)."""
    lib = mesh._lib()

    def stream(tag, m, lo=-1.0, width=2.0):
        return np.array([lo + width * lib.tfs_rng_u01(SEED, tag, i) for i in range(m)])
    na = n_cells + n_bnd
    u, v, w, pp = (stream(0x43560000 + k, na) for k in range(4))
    px, py, pz = (stream(0x43560010 + k, n_cells) for k in range(3))
    sav = stream(0x43560020, n_cells, 0.5, 1.0)
    flux = stream(0x43560030, n_faces)
    _, bc = grad_inputs(n_cells, n_bnd, True)
    return dict(u=u, v=v, w=w, pp=pp, px=px, py=py, pz=pz, sav=sav, flux=flux, bc=bc)


SIMPLE_KINDS = (0, 1, 2, 4, 5, 6)  # wall, symmetry, pressure, inflow, outflow, convective


def simple_kinds(n_bnd):
    """Seeded boundary kinds of the SIMPLE face-loop checks, drawn from
    SIMPLE_KINDS (every kind but buffer)."""
    lib = mesh._lib()
    return np.array([SIMPLE_KINDS[int(6.0 * lib.tfs_rng_u01(SEED, 0x53494D50, i))]
                     for i in range(n_bnd)], np.int8)
