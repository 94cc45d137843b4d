"""Compute_Pressure_Simple's face loop in the oracle (or_pressure_simple),
pinned on the CPU: face fluxes and A % val bitwise against the reference's
own Compute_Pressure_Simple (compiled into oracle/_ref) through digests it
wrote on the committed fixture meshes and live on every generated mesh where
oracle/_ref is built, with mixed boundary kinds (walls, symmetry planes,
pressure, inflow, outflow and convective boundaries); its right-hand side
through the reference's pressure correction PP % n and updated pressure P % n
(the oracle's Bicg, bitwise the reference's, with the routine's cap of 40 on
the oracle's matrix and right-hand side).  Independently of the reference:
conservation (the right-hand side sums to minus the net boundary flux), a
uniform velocity with no pressure terms gives the flux den U . S through
every face that carries one, the matrix is symmetric with the off-diagonals
-A12, and its row sums are the boundary coefficients."""
import hashlib
import pathlib

import numpy as np
import pytest

from mesh_fixtures import CV_DEN, cv_inputs, meta, simple_kinds, unpack
from tflows_amd import mesh

GEN = pathlib.Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "gen"
META = meta()
FIXTURES = list(META["meshes"])
DT = META["dt"]


def sha(a):
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def mesh_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("psmeshes")
    unpack(d)
    return d


def setup(stem):
    stem = str(stem)
    cns = mesh.read_cns(stem + ".cns")
    geo = mesh.read_geo(stem + ".geo", cns)
    scoef, d = mesh.face_coef(cns, geo)
    nc = cns["n_cells"]
    row, col, dia, pos = mesh.topology(nc, cns["faces_c"], base=1, with_pos=True)
    return cns, geo, d, row, col, dia, pos


def run(oracle, cns, geo, d, pos, dia, nnz, cv, bc):
    nc, nb = cns["n_cells"], cns["n_bnd_cells"]
    return oracle.pressure_simple(nc, cns["faces_c"], geo["f"], geo["s"], d, geo["vol"],
                                  cv["sav"], cv["u"], cv["v"], cv["w"], cv["pp"], cv["px"],
                                  cv["py"], cv["pz"], np.full(nc + nb, CV_DEN), bc, pos, dia, nnz)


def correction(oracle, row, col, dia, out, p_old, nc):
    pp = oracle.solve("bicg", row - 1, col - 1, dia - 1, out["val"], out["b"], prec=2, niter=40,
                      tol=1e-9)["x"]
    return pp, p_old[:nc] + 0.8 * pp


@pytest.mark.parametrize("name", FIXTURES)
def test_matches_reference_digests(oracle, mesh_dir, name):
    cns, geo, d, row, col, dia, pos = setup(mesh_dir / name)
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    cv = cv_inputs(nc, nb, nf)
    out = run(oracle, cns, geo, d, pos, dia, len(col), cv, simple_kinds(nb))
    want = META["meshes"][name]["pressure_simple"]
    assert sha(out["flux"]) == want["flux"] and sha(out["val"]) == want["val"]
    pp, p_new = correction(oracle, row, col, dia, out, cv["pp"], nc)
    assert sha(pp) == want["pp"] and sha(p_new) == want["p"]


@pytest.mark.parametrize("name", FIXTURES + ["cylinder", "chan"])
def test_matches_reference_live(reference, oracle, mesh_dir, name):
    stem = mesh_dir / name if name in FIXTURES else GEN / name / name
    if not pathlib.Path(str(stem) + ".geo").exists():
        pytest.skip(f"{stem}.geo not made here (make -C oracle mesh-large)")
    cns, geo, d, row, col, dia, pos = setup(stem)
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    reference.mesh(stem, dt=DT)
    rng = np.random.default_rng(33)
    cv = dict(u=rng.standard_normal(nc + nb), v=rng.standard_normal(nc + nb),
              w=rng.standard_normal(nc + nb), pp=rng.standard_normal(nc + nb),
              px=rng.standard_normal(nc), py=rng.standard_normal(nc), pz=rng.standard_normal(nc),
              sav=rng.uniform(0.5, 2.0, nc))
    bc = rng.choice(np.array([0, 1, 2, 4, 5, 6], np.int8), nb)
    ref = reference.mesh_press_simple(nc, nf, len(col), cv["u"], cv["v"], cv["w"], cv["pp"],
                                      cv["px"], cv["py"], cv["pz"], cv["sav"], CV_DEN, bc, 2,
                                      1e-9, 0.8)
    out = run(oracle, cns, geo, d, pos, dia, len(col), cv, bc)
    assert np.array_equal(out["flux"], ref["flux"]) and np.array_equal(out["val"], ref["val"])
    pp, p_new = correction(oracle, row, col, dia, out, cv["pp"], nc)
    assert np.array_equal(pp, ref["pp"]) and np.array_equal(p_new, ref["p"])


@pytest.mark.parametrize("name", FIXTURES)
def test_conservation_and_structure(oracle, mesh_dir, name):
    cns, geo, d, row, col, dia, pos = setup(mesh_dir / name)
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    cv = cv_inputs(nc, nb, nf)
    bc = simple_kinds(nb)
    out = run(oracle, cns, geo, d, pos, dia, len(col), cv, bc)
    fc = cns["faces_c"]
    c2 = fc[1::2]
    bnd = c2 < 0
    kind = np.full(nf, -1)
    kind[bnd] = bc[-c2[bnd] - 1]
    # conservation: interior fluxes cancel in b
    assert abs(out["b"].sum() + out["flux"][bnd].sum()) <= 1e-10 * np.abs(out["flux"]).sum()
    # walls and symmetry planes carry no flux and add nothing to the matrix
    none = bnd & np.isin(kind, (0, 1))
    assert (out["flux"][none] == 0).all() and (out["a12"][none] == 0).all()
    assert (out["bnd_on"] == (bnd & np.isin(kind, (2, 5, 6)))).all()
    assert (out["a12"][~bnd] > 0).all()
    # the matrix: symmetric, off-diagonals -A12, row sums the boundary A12
    val = out["val"]
    assert np.array_equal(val[pos[0::2][~bnd] - 1], -out["a12"][~bnd])
    assert np.array_equal(val[pos[1::2][~bnd] - 1], -out["a12"][~bnd])
    rows = np.repeat(np.arange(nc), np.diff(row))
    rs = np.zeros(nc)
    np.add.at(rs, rows, val)
    rb = np.zeros(nc)
    np.add.at(rb, fc[0::2][bnd] - 1, out["a12"][bnd])
    assert np.allclose(rs, rb, rtol=0, atol=1e-12 * np.abs(val).max())
    # a uniform velocity and no pressure terms: den U . S wherever a flux is carried
    u0 = np.array([0.3, -0.2, 0.5])
    ones, zc = np.ones(nc + nb), np.zeros(nc)
    cu = dict(u=u0[0] * ones, v=u0[1] * ones, w=u0[2] * ones, pp=np.zeros(nc + nb),
              px=zc, py=zc, pz=zc, sav=cv["sav"])
    flux = run(oracle, cns, geo, d, pos, dia, len(col), cu, bc)["flux"]
    want = np.where(none, 0.0, CV_DEN * (u0 @ geo["s"]))
    assert np.abs(flux - want).max() <= 1e-12 * np.abs(want).max()
