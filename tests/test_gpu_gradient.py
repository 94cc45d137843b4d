"""Least-squares gradients on the GPU (tfs_gradient_*: Compute_Gradient_Matrix,
GraPhi, GradP) bitwise against the oracle -- itself bitwise equal to the
reference's routines -- on the reference generator's meshes, for walls and
for mixed wall / symmetry / pressure boundaries; the reference's own digests
of the same computation; a linear field's gradient; and the refusals
(a boundary cell on two faces, an unknown boundary kind, GradP without
cell centres)."""
import hashlib

import numpy as np
import pytest

from mesh_fixtures import grad_inputs, meta, unpack
from tflows_amd import mesh

pytestmark = pytest.mark.gpu

META = meta()
NAMES = list(META["meshes"])


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.fail("no GPU visible")
    return t


@pytest.fixture(scope="module")
def S():
    from tflows_amd import solvers
    return solvers


@pytest.fixture(scope="module")
def meshes(tmp_path_factory):
    d = tmp_path_factory.mktemp("gradgpu")
    out = {}
    for name in unpack(d):
        cns = mesh.read_cns(d / (name + ".cns"))
        geo = mesh.read_geo(d / (name + ".geo"), cns)
        _, dv = mesh.face_coef(cns, geo)
        out[name] = (cns, geo, dv)
    return out


def sha(a):
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()).hexdigest()


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("mixed", [False, True])
def test_gradients_bitwise(torch, S, oracle, meshes, name, mixed):
    cns, geo, d = meshes[name]
    nc, nb, faces = cns["n_cells"], cns["n_bnd_cells"], cns["faces_c"]
    print(f"gradients {name} n={nc} mixed={mixed}", flush=True)
    phi, bc = grad_inputs(nc, nb, mixed)
    G = S.Gradient(nc, nb, faces, d, geo["xc"], bc)
    g_or = oracle.gradient_matrix(nc, faces, d)
    assert np.array_equal(G.matrix(), g_or)
    want = META["meshes"][name]["gradients"]
    assert sha(G.matrix()) == want["G"]
    tag = "mixed" if mixed else "walls"
    dphi = torch.from_numpy(phi).cuda()
    p = [torch.zeros(nc, dtype=torch.float64, device="cuda") for _ in range(3)]
    G.grad_phi(dphi, *p)
    got = [t.cpu().numpy() for t in p]
    for a, b in zip(got, oracle.graphi(nc, faces, d, g_or, phi, bc)):
        assert np.array_equal(a, b)
    assert [sha(a) for a in got] == want[f"graphi_{tag}"]
    G.grad_p(dphi, *p)
    got = [dphi.cpu().numpy()] + [t.cpu().numpy() for t in p]
    for a, b in zip(got, oracle.gradp(nc, faces, d, g_or, geo["xc"], phi, bc)):
        assert np.array_equal(a, b)
    assert [sha(a) for a in got] == want[f"gradp_{tag}"]


def test_linear_field(torch, S, meshes):
    cns, geo, d = meshes["backstep_nonorthogonal"]
    nc, nb = cns["n_cells"], cns["n_bnd_cells"]
    G = S.Gradient(nc, nb, cns["faces_c"], d)
    a = np.array([-0.4, 1.9, 0.25])
    phi = torch.from_numpy(a @ geo["xc"]).cuda()
    p = [torch.zeros(nc, dtype=torch.float64, device="cuda") for _ in range(3)]
    G.grad_phi(phi, *p)
    f = cns["faces_c"].reshape(-1, 2)
    per = np.abs(geo["shift"]).sum(0) > 0
    skip = np.zeros(nc, bool)
    skip[f[per, 0] - 1] = True
    skip[f[per & (f[:, 1] > 0), 1] - 1] = True
    for k in range(3):
        assert np.abs(p[k].cpu().numpy()[~skip] - a[k]).max() < 1e-9 * np.abs(a).max()


def test_refusals(torch, S, meshes):
    cns, geo, d = meshes["cavity"]
    nc, nb = cns["n_cells"], cns["n_bnd_cells"]
    faces = cns["faces_c"].copy()
    bnd = np.flatnonzero(faces[1::2] < 0)
    faces[2 * bnd[1] + 1] = faces[2 * bnd[0] + 1]        # one boundary cell on two faces
    with pytest.raises(S.SolverError):
        S.Gradient(nc, nb, faces, d)
    with pytest.raises(S.SolverError):
        S.Gradient(nc, nb, cns["faces_c"], d, bc=np.full(nb, 7, np.uint8))
    G = S.Gradient(nc, nb, cns["faces_c"], d)            # no centres: GradP refused
    phi = torch.zeros(nc + nb, dtype=torch.float64, device="cuda")
    p = [torch.zeros(nc, dtype=torch.float64, device="cuda") for _ in range(3)]
    with pytest.raises(S.SolverError):
        G.grad_p(phi, *p)


@pytest.mark.parametrize("name", NAMES)
def test_correct_velocity_bitwise(torch, S, oracle, meshes, name):
    """Correct_Velocity on the GPU against the oracle (every array bitwise,
    the three maxima exactly) and the reference's digests."""
    from mesh_fixtures import CV_DEN, CV_VISC, cv_inputs
    cns, geo, d = meshes[name]
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    print(f"correct_velocity {name} n={nc}", flush=True)
    dt = META["dt"]
    scoef, _ = mesh.face_coef(cns, geo)
    row, col, dia, pos = mesh.topology(nc, cns["faces_c"], base=1, with_pos=True)
    val = mesh.pressure_values(nc, cns["faces_c"], pos, dia, len(col), scoef, dt, base=1)
    coef = np.where(pos[0::2] > 0, val[np.maximum(pos[0::2], 1) - 1], 0.0)
    cv = cv_inputs(nc, nb, nf)
    den = np.full(nc, CV_DEN)
    want = oracle.correct_velocity(nc, cns["faces_c"], geo["vol"], cv["sav"], cv["px"], cv["py"],
                                   cv["pz"], cv["u"], cv["v"], cv["w"], cv["pp"], coef, scoef, d,
                                   dt, CV_VISC, cv["flux"], den, cv["bc"])
    G = S.Gradient(nc, nb, cns["faces_c"], d, geo["xc"], cv["bc"])

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    u, v, w, flux = dev(cv["u"]), dev(cv["v"]), dev(cv["w"]), dev(cv["flux"])
    b = torch.zeros(nc, dtype=torch.float64, device="cuda")
    em, cm, pm = G.correct_velocity(dev(geo["vol"]), dev(cv["sav"]), dev(cv["px"]),
                                    dev(cv["py"]), dev(cv["pz"]), u, v, w, dev(cv["pp"]),
                                    dev(coef), dev(den), dev(scoef), dt, CV_VISC, flux, b)
    got = dict(u=u, v=v, w=w, flux=flux, b=b)
    ref = META["meshes"][name]["correct_velocity"]
    for k, t in got.items():
        a = t.cpu().numpy()
        assert np.array_equal(a, want[k]), k
        assert sha(a) == ref[k], k
    assert em == want["errmax"] == float.fromhex(ref["errmax"])
    assert cm == want["cfl"] and pm == want["pe"]


@pytest.mark.parametrize("name", NAMES)
def test_pressure_fractional_bitwise(torch, S, oracle, meshes, name):
    """Compute_Pressure_Fractional's face loop on the GPU (fluxes, A12, the
    right-hand side) and the matrix assembled from its A12 on the GPU,
    against the oracle bitwise and the reference's digests."""
    from mesh_fixtures import CV_DEN, cv_inputs
    cns, geo, d = meshes[name]
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    print(f"pressure_fractional {name} n={nc}", flush=True)
    cv = cv_inputs(nc, nb, nf)
    den = np.full(nc + nb, CV_DEN)
    flux_or, a12_or, b_or = oracle.pressure_fractional(
        nc, cns["faces_c"], geo["f"], geo["s"], geo["vol"], cv["sav"], cv["u"], cv["v"], cv["w"],
        cv["px"], cv["py"], cv["pz"], cv["pp"], den)
    G = S.Gradient(nc, nb, cns["faces_c"], d)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float64).ravel()).cuda()
    flux = torch.zeros(nf, dtype=torch.float64, device="cuda")
    a12 = torch.zeros(nf, dtype=torch.float64, device="cuda")
    b = torch.zeros(nc, dtype=torch.float64, device="cuda")
    G.pressure_fractional(dev(geo["f"]), dev(geo["s"]), dev(geo["vol"]), dev(cv["sav"]),
                          dev(cv["u"]), dev(cv["v"]), dev(cv["w"]), dev(cv["px"]), dev(cv["py"]),
                          dev(cv["pz"]), dev(cv["pp"]), dev(den), flux, a12, b)
    assert np.array_equal(flux.cpu().numpy(), flux_or)
    assert np.array_equal(a12.cpu().numpy(), a12_or)
    assert np.array_equal(b.cpu().numpy(), b_or)
    row, col, dia, pos = mesh.topology(nc, cns["faces_c"], base=1, with_pos=True)
    A = S.Matrix(row - 1, col - 1, dia - 1)
    S.Assembly(A, cns["faces_c"], pos, dia, base=1).assemble(a12)
    val = A.get_values()
    assert np.array_equal(val, oracle.assemble(nc, cns["faces_c"], pos, dia, len(col), a12_or))
    want = META["meshes"][name]["pressure_fractional"]
    assert sha(flux.cpu().numpy()) == want["flux"] and sha(val) == want["val"]


@pytest.mark.parametrize("name", NAMES)
def test_pressure_simple_bitwise(torch, S, oracle, meshes, name):
    """Compute_Pressure_Simple's face loop on the GPU (fluxes, A12, which
    boundary faces carry a coefficient, the right-hand side), with mixed
    boundary kinds, and the matrix assembled from its A12 on the GPU, against
    the oracle bitwise and the reference's digests."""
    from mesh_fixtures import CV_DEN, cv_inputs, simple_kinds
    cns, geo, d = meshes[name]
    nc, nb, nf = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"]
    print(f"pressure_simple {name} n={nc}", flush=True)
    cv = cv_inputs(nc, nb, nf)
    bc = simple_kinds(nb)
    den = np.full(nc + nb, CV_DEN)
    row, col, dia, pos = mesh.topology(nc, cns["faces_c"], base=1, with_pos=True)
    want = oracle.pressure_simple(nc, cns["faces_c"], geo["f"], geo["s"], d, geo["vol"],
                                  cv["sav"], cv["u"], cv["v"], cv["w"], cv["pp"], cv["px"],
                                  cv["py"], cv["pz"], den, bc, pos, dia, len(col))
    G = S.Gradient(nc, nb, cns["faces_c"], d, bc=bc)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float64).ravel()).cuda()
    flux = torch.full((nf,), np.nan, dtype=torch.float64, device="cuda")
    a12 = torch.full((nf,), np.nan, dtype=torch.float64, device="cuda")
    bnd_on = torch.full((nf,), 7, dtype=torch.uint8, device="cuda")
    b = torch.full((nc,), np.nan, dtype=torch.float64, device="cuda")
    G.pressure_simple(dev(geo["f"]), dev(geo["s"]), dev(geo["vol"]), dev(cv["sav"]),
                      dev(cv["u"]), dev(cv["v"]), dev(cv["w"]), dev(cv["px"]), dev(cv["py"]),
                      dev(cv["pz"]), dev(cv["pp"]), dev(den), flux, a12, bnd_on, b)
    assert np.array_equal(flux.cpu().numpy(), want["flux"])
    assert np.array_equal(a12.cpu().numpy(), want["a12"])
    assert np.array_equal(bnd_on.cpu().numpy(), want["bnd_on"])
    assert np.array_equal(b.cpu().numpy(), want["b"])
    A = S.Matrix(row - 1, col - 1, dia - 1)
    S.Assembly(A, cns["faces_c"], pos, dia, base=1).assemble(a12, bnd_on=bnd_on)
    val = A.get_values()
    assert np.array_equal(val, want["val"])
    ref = META["meshes"][name]["pressure_simple"]
    assert sha(flux.cpu().numpy()) == ref["flux"] and sha(val) == ref["val"]
