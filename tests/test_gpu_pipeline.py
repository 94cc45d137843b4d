"""One fractional-step pressure correction through the device entry points
on a real T-FlowS mesh (the reference generator's non-orthogonal backstep),
against the same steps through the oracle on the CPU:

  faces -> tfs_topology_build_device -> tfs_assemble_device (dt Scoef, walls)
  -> cell 1 pinned -> Cg + DIC (prec 2) to 1e-10 ||b|| -> tfs_grad_p_device
  -> tfs_correct_velocity_device.

Topology and assembly are bitwise; the solve is within the solve bar
(rel-L2 1e-8); what follows it is linear in the solution, so it is checked
to 1e-6 of each field's scale."""
import numpy as np
import pytest

from conftest import rel_l2
from mesh_fixtures import CV_DEN, CV_VISC, cv_inputs, meta, unpack
from tflows_amd import mesh

pytestmark = pytest.mark.gpu

DT = meta()["dt"]


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.fail("no GPU visible")
    return t


def test_pressure_correction_pipeline(torch, oracle, tmp_path):
    from tflows_amd import solvers as S
    unpack(tmp_path)
    stem = str(tmp_path / "backstep_nonorthogonal")
    cns = mesh.read_cns(stem + ".cns")
    geo = mesh.read_geo(stem + ".geo", cns)
    scoef, d = mesh.face_coef(cns, geo)
    nc, nb, nf, faces = cns["n_cells"], cns["n_bnd_cells"], cns["n_faces"], cns["faces_c"]
    print(f"pipeline backstep n={nc}", flush=True)

    def dev(a, dtype=None):
        return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)

    # topology on the device, bitwise the host's
    row, col, dia, pos = (t.cpu().numpy() for t in S.topology_device(nc, dev(faces), base=1))
    want = mesh.topology(nc, faces, base=1, with_pos=True)
    for a, b in zip((row, col, dia, pos), want):
        assert np.array_equal(a, b)
    # pressure matrix assembled on the device (walls add nothing), then pinned
    A = S.Matrix(row - 1, col - 1, dia - 1)
    asm = S.Assembly(A, faces, pos, dia, base=1)
    a12 = DT * scoef
    a12[faces[1::2] < 0] = 0.0
    asm.assemble(dev(a12))
    val = A.get_values()
    assert np.array_equal(val, mesh.pressure_values(nc, faces, pos, dia, len(col), scoef, DT,
                                                    base=1))
    val[dia[0] - 1] *= 2.0
    A.set_values(val)
    # the solve
    c = mesh.mesh_case(stem, dt=DT)
    tol = 1e-10 * float(np.linalg.norm(c.b))
    x = torch.zeros(nc, dtype=torch.float64, device="cuda")
    niter, ini, fin = S.Cg(A, x, dev(c.b), 2, 20000, tol)
    ref = oracle.solve("cg", c.row, c.col, c.dia, c.val, c.b, prec=2, niter=20000, tol=tol)
    assert fin < tol and abs(niter - ref["niter"]) <= max(1, ref["niter"] // 20)
    assert rel_l2(x.cpu().numpy(), ref["x"]) <= 1e-8
    # pressure gradient (GradP) of the solution, boundary cells extrapolated
    cv = cv_inputs(nc, nb, nf)
    G = S.Gradient(nc, nb, faces, d, geo["xc"], cv["bc"])
    pp = np.concatenate([ref["x"], np.zeros(nb)])
    dpp = torch.cat([x, torch.zeros(nb, dtype=torch.float64, device="cuda")])
    grads = [torch.zeros(nc, dtype=torch.float64, device="cuda") for _ in range(3)]
    G.grad_p(dpp, *grads)
    g_or = oracle.gradient_matrix(nc, faces, d)
    pp_or, px, py, pz = oracle.gradp(nc, faces, d, g_or, geo["xc"], pp, cv["bc"])
    for got, want in zip([dpp] + grads, (pp_or, px, py, pz)):
        got = got.cpu().numpy()
        assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()
    # velocity correction with the pressure correction just computed
    coef = np.where(pos[0::2] > 0, val[np.maximum(pos[0::2], 1) - 1], 0.0)
    den = np.full(nc, CV_DEN)
    out = oracle.correct_velocity(nc, faces, geo["vol"], cv["sav"], px, py, pz, cv["u"], cv["v"],
                                  cv["w"], pp_or, coef, scoef, d, DT, CV_VISC, cv["flux"], den,
                                  cv["bc"])
    u, v, w, flux = dev(cv["u"]), dev(cv["v"]), dev(cv["w"]), dev(cv["flux"])
    b = torch.zeros(nc, dtype=torch.float64, device="cuda")
    em, cm, pm = G.correct_velocity(dev(geo["vol"]), dev(cv["sav"]), *grads, u, v, w, dpp,
                                    dev(coef), dev(den), dev(scoef), DT, CV_VISC, flux, b)
    for k, t in (("u", u), ("v", v), ("w", w), ("flux", flux), ("b", b)):
        got = t.cpu().numpy()
        assert np.abs(got - out[k]).max() <= 1e-6 * max(np.abs(out[k]).max(), 1.0), k
    assert abs(em - out["errmax"]) <= 1e-6 * out["errmax"]
