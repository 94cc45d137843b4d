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


# ---- Compute_Pressure_Fractional's face loop ----------------------------------

def _pf(oracle, cns, geo, cv):
    nc, nb = cns["n_cells"], cns["n_bnd_cells"]
    return oracle.pressure_fractional(nc, cns["faces_c"], geo["f"], geo["s"], geo["vol"],
                                      cv["sav"], cv["u"], cv["v"], cv["w"], cv["px"], cv["py"],
                                      cv["pz"], cv["pp"], np.full(nc + nb, CV_DEN))


@pytest.mark.parametrize("name", FIXTURES)
def test_pressure_fractional_matches_reference_digests(oracle, mesh_dir, name):
    """Fluxes and matrix bitwise; the right-hand side through the reference's
    own pressure-correction solution (the oracle's Cg, bitwise the
    reference's, on the oracle's matrix and right-hand side)."""
    cns, geo, scoef, d, coef = setup(mesh_dir / name)
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    cv = cv_inputs(nc, nb, nf)
    flux, a12, b = _pf(oracle, cns, geo, cv)
    row, col, dia, pos = mesh.topology(nc, cns["faces_c"], base=1, with_pos=True)
    val = oracle.assemble(nc, cns["faces_c"], pos, dia, len(col), a12)
    want = META["meshes"][name]["pressure_fractional"]
    assert sha(flux) == want["flux"] and sha(val) == want["val"]
    pp = oracle.solve("cg", row - 1, col - 1, dia - 1, val, b, prec=2, niter=200, tol=1e-9)["x"]
    assert sha(pp) == want["pp"]
    p_new = cv["pp"][:nc] + 0.8 * pp
    p_new = p_new - 0.5 * (p_new.max() + p_new.min())
    assert sha(p_new) == want["p"]


@pytest.mark.parametrize("name", FIXTURES)
def test_pressure_fractional_conservation(oracle, mesh_dir, name):
    cns, geo, scoef, d, coef = setup(mesh_dir / name)
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    cv = cv_inputs(nc, nb, nf)
    flux, a12, b = _pf(oracle, cns, geo, cv)
    c2 = cns["faces_c"][1::2]
    assert abs(b.sum() + flux[c2 < 0].sum()) <= 1e-10 * np.abs(flux).sum()
    assert (a12[c2 > 0] > 0).all() and (a12[c2 < 0] == 0).all()
    # a uniform field u = U0 with no pressure terms: the flux is den U0 . S
    zero = np.zeros(nc)
    u0 = np.array([0.3, -0.2, 0.5])
    ones = np.ones(nc + nb)
    flux, _, _ = oracle.pressure_fractional(nc, cns["faces_c"], geo["f"], geo["s"], geo["vol"],
                                            cv["sav"], u0[0] * ones, u0[1] * ones, u0[2] * ones,
                                            zero, zero, zero, np.zeros(nc + nb),
                                            np.full(nc + nb, CV_DEN))
    want = CV_DEN * (u0 @ geo["s"])
    assert np.abs(flux - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("name", FIXTURES + ["cylinder", "chan"])
def test_pressure_fractional_matches_reference_live(reference, oracle, mesh_dir, name):
    stem = mesh_dir / name if name in FIXTURES else GEN / name / name
    if not pathlib.Path(str(stem) + ".geo").exists():
        pytest.skip(f"{stem}.geo not made here (make -C oracle mesh-large)")
    cns, geo, scoef, d, coef = setup(stem)
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    row, col, dia, pos = mesh.topology(nc, cns["faces_c"], base=1, with_pos=True)
    reference.mesh(stem, dt=DT)
    rng = np.random.default_rng(21)
    cv = dict(u=rng.standard_normal(nc + nb), v=rng.standard_normal(nc + nb),
              w=rng.standard_normal(nc + nb), pp=rng.standard_normal(nc + nb),
              px=rng.standard_normal(nc), py=rng.standard_normal(nc), pz=rng.standard_normal(nc),
              sav=rng.uniform(0.5, 2.0, nc))
    ref = reference.mesh_press_frac(nc, nf, len(col), cv["u"], cv["v"], cv["w"], cv["pp"],
                                    cv["px"], cv["py"], cv["pz"], cv["sav"], CV_DEN, 2, 1e-9, 0.8)
    flux, a12, b = _pf(oracle, cns, geo, cv)
    val = oracle.assemble(nc, cns["faces_c"], pos, dia, len(col), a12)
    assert np.array_equal(flux, ref["flux"]) and np.array_equal(val, ref["val"])
    pp = oracle.solve("cg", row - 1, col - 1, dia - 1, val, b, prec=2, niter=200, tol=1e-9)["x"]
    assert np.array_equal(pp, ref["pp"])
