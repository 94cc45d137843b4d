"""The least-squares gradients of the oracle (or_gradient_matrix, or_graphi,
or_gradp: Compute_Gradient_Matrix, GraPhi, GradP), pinned on the CPU:

  * bitwise against the reference's own routines on the committed fixture
    meshes, through SHA-256 digests written by the reference
    (tests/golden/make_mesh_golden.py), for walls and for mixed wall /
    symmetry / pressure boundaries;
  * live against the reference compiled into oracle/_ref where it was built,
    also on its larger refined and periodic meshes;
  * against what the mathematics fixes: G is the inverse of the cell's normal
    matrix sum d d^T; a linear field's gradient is exact in every cell whose
    faces cross no periodic boundary (least squares reproduces linear data);
    a symmetry boundary gives the gradient of the field with that boundary
    cell set to its inside value; GradP leaves pressure boundaries alone and
    on a linear field changes only cells next to copied boundaries.
"""
import hashlib
import pathlib

import numpy as np
import pytest

from mesh_fixtures import grad_inputs, meta, unpack
from tflows_amd import mesh

GEN = pathlib.Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "gen"
META = meta()
FIXTURES = list(META["meshes"])


def sha(a):
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def mesh_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("gradmeshes")
    unpack(d)
    return d


_cache = {}


def geometry(stem):
    stem = str(stem)
    if stem not in _cache:
        cns = mesh.read_cns(stem + ".cns")
        geo = mesh.read_geo(stem + ".geo", cns)
        _, d = mesh.face_coef(cns, geo)
        _cache[stem] = (cns, geo, d)
    return _cache[stem]


@pytest.mark.parametrize("name", FIXTURES)
def test_matches_reference_digests(oracle, mesh_dir, name):
    cns, geo, d = geometry(mesh_dir / name)
    nc, nb, faces = cns["n_cells"], cns["n_bnd_cells"], cns["faces_c"]
    want = META["meshes"][name]["gradients"]
    g = oracle.gradient_matrix(nc, faces, d)
    assert sha(g) == want["G"]
    for tag, mixed in (("walls", False), ("mixed", True)):
        phi, bc = grad_inputs(nc, nb, mixed)
        got = oracle.graphi(nc, faces, d, g, phi, bc)
        assert [sha(a) for a in got] == want[f"graphi_{tag}"], tag
        got = oracle.gradp(nc, faces, d, g, geo["xc"], phi, bc)
        assert [sha(a) for a in got] == want[f"gradp_{tag}"], tag


@pytest.mark.parametrize("name", FIXTURES + ["cylinder", "chan"])
def test_matches_reference_live(reference, oracle, mesh_dir, name):
    stem = mesh_dir / name if name in FIXTURES else GEN / name / name
    if not pathlib.Path(str(stem) + ".geo").exists():
        pytest.skip(f"{stem}.geo not made here (make -C oracle mesh-large)")
    cns, geo, d = geometry(stem)
    nc, nb, faces = cns["n_cells"], cns["n_bnd_cells"], cns["faces_c"]
    reference.mesh(stem)
    g = oracle.gradient_matrix(nc, faces, d)
    assert np.array_equal(reference.mesh_grad(0, nc, nb), g)
    rng = np.random.default_rng(8)
    phi = rng.standard_normal(nc + nb)
    bc = rng.integers(0, 3, nb).astype(np.int8)
    for a, b in zip(reference.mesh_grad(1, nc, nb, phi, bc), oracle.graphi(nc, faces, d, g, phi, bc)):
        assert np.array_equal(a, b)
    for a, b in zip(reference.mesh_grad(2, nc, nb, phi, bc),
                    oracle.gradp(nc, faces, d, g, geo["xc"], phi, bc)):
        assert np.array_equal(a, b)


def _periodic_cells(cns, geo):
    f = cns["faces_c"].reshape(-1, 2)
    per = np.abs(geo["shift"]).sum(0) > 0
    mark = np.zeros(cns["n_cells"], bool)
    mark[f[per, 0] - 1] = True
    inn = per & (f[:, 1] > 0)
    mark[f[inn, 1] - 1] = True
    return mark


@pytest.mark.parametrize("name", FIXTURES)
def test_gradient_matrix_is_inverse_normal_matrix(oracle, mesh_dir, name):
    cns, geo, d = geometry(mesh_dir / name)
    nc, faces = cns["n_cells"], cns["faces_c"].reshape(-1, 2)
    g = oracle.gradient_matrix(nc, cns["faces_c"], d)
    m = np.zeros((nc, 3, 3))
    dd = d.T[:, :, None] * d.T[:, None, :]
    np.add.at(m, faces[:, 0] - 1, dd)
    inn = faces[:, 1] > 0
    np.add.at(m, faces[inn, 1] - 1, dd[inn])
    G = g[:, [0, 3, 4, 3, 1, 5, 4, 5, 2]].reshape(nc, 3, 3)
    prod = np.einsum("cij,cjk->cik", G, m)
    assert np.abs(prod - np.eye(3)).max() < 1e-9


@pytest.mark.parametrize("name", FIXTURES)
def test_linear_field_exact(oracle, mesh_dir, name):
    cns, geo, d = geometry(mesh_dir / name)
    nc = cns["n_cells"]
    g = oracle.gradient_matrix(nc, cns["faces_c"], d)
    a = np.array([0.7, -1.3, 2.1])
    phi = a @ geo["xc"]                 # on cells and boundary cells
    px, py, pz = oracle.graphi(nc, cns["faces_c"], d, g, phi)
    ok = ~_periodic_cells(cns, geo)
    assert ok.sum() > nc // 4
    for k, p in enumerate((px, py, pz)):
        assert np.abs(p[ok] - a[k]).max() < 1e-9 * np.abs(a).max()


@pytest.mark.parametrize("name", FIXTURES)
def test_symmetry_equals_inside_value(oracle, mesh_dir, name):
    cns, geo, d = geometry(mesh_dir / name)
    nc, nb, faces = cns["n_cells"], cns["n_bnd_cells"], cns["faces_c"].reshape(-1, 2)
    g = oracle.gradient_matrix(nc, cns["faces_c"], d)
    phi, _ = grad_inputs(nc, nb, False)
    sym = np.ones(nb, np.int8)
    got = oracle.graphi(nc, cns["faces_c"], d, g, phi, sym)
    copied = phi.copy()
    bnd = faces[:, 1] < 0
    copied[nc + (-faces[bnd, 1] - 1)] = phi[faces[bnd, 0] - 1]
    want = oracle.graphi(nc, cns["faces_c"], d, g, copied)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", FIXTURES)
def test_gradp_boundary_rules(oracle, mesh_dir, name):
    cns, geo, d = geometry(mesh_dir / name)
    nc, nb = cns["n_cells"], cns["n_bnd_cells"]
    faces = cns["faces_c"].reshape(-1, 2)
    g = oracle.gradient_matrix(nc, cns["faces_c"], d)
    phi, bc = grad_inputs(nc, nb, True)
    out, px, py, pz = oracle.gradp(nc, cns["faces_c"], d, g, geo["xc"], phi, bc)
    assert np.array_equal(out[:nc], phi[:nc])              # cells untouched
    press = np.flatnonzero(bc == 2)
    assert np.array_equal(out[nc + press], phi[nc + press])  # pressure boundaries untouched
    # the other boundary cells: inside value + 1.2 grad . (x_b - x_c)
    bnd = faces[:, 1] < 0
    c1, b = faces[bnd, 0] - 1, -faces[bnd, 1] - 1
    keep = bc[b] != 2
    c1, b = c1[keep], b[keep]
    xc = geo["xc"]
    grad1 = oracle.graphi(nc, cns["faces_c"], d, g, np.where(
        np.isin(np.arange(nc + nb), nc + b), 0.0, phi) + np.bincount(
        nc + b, weights=phi[c1], minlength=nc + nb), bc)
    ext = phi[c1] + 1.2 * (grad1[0][c1] * (xc[0, nc + b] - xc[0, c1]) +
                           grad1[1][c1] * (xc[1, nc + b] - xc[1, c1]) +
                           grad1[2][c1] * (xc[2, nc + b] - xc[2, c1]))
    assert np.array_equal(out[nc + b], ext)
