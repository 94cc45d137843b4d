"""Correct_Velocity in the oracle (or_correct_velocity), pinned on the CPU:
bitwise against the reference's own Correct_Velocity (velocities, fluxes,
mass errors, errmax) on the committed fixture meshes through digests it
wrote, live on all generated meshes where oracle/_ref is built, and the CFL /
Peclet maxima to the four digits the reference prints; plus what the
discrete conservation fixes: the volume-weighted mass errors sum to minus
the net boundary outflow (interior fluxes cancel pairwise), and a uniform
pressure correction leaves every flux unchanged."""
import hashlib
import pathlib

import numpy as np
import pytest

from mesh_fixtures import CV_DEN, CV_VISC, cv_inputs, meta, unpack
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
    d = tmp_path_factory.mktemp("cvmeshes")
    unpack(d)
    return d


def setup(stem):
    stem = str(stem)
    cns = mesh.read_cns(stem + ".cns")
    geo = mesh.read_geo(stem + ".geo", cns)
    scoef, d = mesh.face_coef(cns, geo)
    nc = cns["n_cells"]
    row, col, dia, pos = mesh.topology(nc, cns["faces_c"], base=1, with_pos=True)
    val = mesh.pressure_values(nc, cns["faces_c"], pos, dia, len(col), scoef, DT, base=1)
    coef = np.where(pos[0::2] > 0, val[np.maximum(pos[0::2], 1) - 1], 0.0)
    return cns, geo, scoef, d, coef


def run(oracle, cns, geo, scoef, d, coef, cv):
    nc = cns["n_cells"]
    return oracle.correct_velocity(nc, cns["faces_c"], geo["vol"], cv["sav"], cv["px"], cv["py"],
                                   cv["pz"], cv["u"], cv["v"], cv["w"], cv["pp"], coef, scoef, d,
                                   DT, CV_VISC, cv["flux"], np.full(nc, CV_DEN), cv["bc"])


@pytest.mark.parametrize("name", FIXTURES)
def test_matches_reference_digests(oracle, mesh_dir, name):
    cns, geo, scoef, d, coef = setup(mesh_dir / name)
    cv = cv_inputs(cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"])
    out = run(oracle, cns, geo, scoef, d, coef, cv)
    want = META["meshes"][name]["correct_velocity"]
    for k in ("u", "v", "w", "flux", "b"):
        assert sha(out[k]) == want[k], k
    assert out["errmax"] == float.fromhex(want["errmax"])
    assert float(f"{out['cfl']:.3e}") == want["cfl"]
    assert float(f"{out['pe']:.3e}") == want["pe"]


@pytest.mark.parametrize("name", FIXTURES + ["cylinder", "chan"])
def test_matches_reference_live(reference, oracle, mesh_dir, name):
    stem = mesh_dir / name if name in FIXTURES else GEN / name / name
    if not pathlib.Path(str(stem) + ".geo").exists():
        pytest.skip(f"{stem}.geo not made here (make -C oracle mesh-large)")
    cns, geo, scoef, d, coef = setup(stem)
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    reference.mesh(stem, dt=DT)
    rng = np.random.default_rng(12)
    cv = dict(u=rng.standard_normal(nc + nb), v=rng.standard_normal(nc + nb),
              w=rng.standard_normal(nc + nb), pp=rng.standard_normal(nc + nb),
              px=rng.standard_normal(nc), py=rng.standard_normal(nc), pz=rng.standard_normal(nc),
              sav=rng.uniform(0.5, 2.0, nc), flux=rng.standard_normal(nf),
              bc=rng.integers(0, 3, nb).astype(np.int8))
    ref = reference.mesh_correct(nc, cv["u"], cv["v"], cv["w"], cv["px"], cv["py"], cv["pz"],
                                 cv["pp"], cv["sav"], cv["flux"], CV_DEN, CV_VISC, cv["bc"])
    out = run(oracle, cns, geo, scoef, d, coef, cv)
    for k in ("u", "v", "w", "flux", "b"):
        assert np.array_equal(out[k], ref[k]), k
    assert out["errmax"] == ref["errmax"]
    assert float(f"{out['cfl']:.3e}") == ref["cfl"] and float(f"{out['pe']:.3e}") == ref["pe"]


@pytest.mark.parametrize("name", FIXTURES)
def test_conservation(oracle, mesh_dir, name):
    cns, geo, scoef, d, coef = setup(mesh_dir / name)
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    cv = cv_inputs(nc, nb, nf)
    out = run(oracle, cns, geo, scoef, d, coef, cv)
    c2 = cns["faces_c"][1::2]
    net = out["b"] @ (geo["vol"] * CV_DEN)
    assert abs(net + out["flux"][c2 < 0].sum()) <= 1e-10 * np.abs(out["flux"]).sum()
    cv["pp"] = np.full(nc + nb, 0.37)
    out = run(oracle, cns, geo, scoef, d, coef, cv)
    assert np.array_equal(out["flux"], cv["flux"])
