"""Matrix_Mod_Topology on the device (tfs_topology_build_device) against the
host builder (libtfs_mesh.so: tfs_topology_build, itself pinned bitwise
against the reference's Matrix_Mod_Topology on the CPU): bitwise equal row,
col, dia and pos on synthetic boxes, permuted and jittered meshes, the
reference generator's meshes, and the edge cases the loops allow (no faces,
only boundary faces, two faces between one pair of cells, a face of a cell
with itself); malformed faces and a too-small column capacity are refused.
"""
import ctypes
import gzip
import pathlib
import shutil

import numpy as np
import pytest

from tflows_amd import mesh

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "mesh"


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


def check(torch, S, n, faces, base):
    faces = np.ascontiguousarray(faces, np.int32)
    want = mesh.topology(n, faces, base=base, with_pos=True)
    got = S.topology_device(n, torch.from_numpy(faces).cuda(), base=base)
    for name, w, g in zip(("row", "col", "dia", "pos"), want, got):
        assert np.array_equal(g.cpu().numpy(), w), name


SYNTH = {
    "box": dict(kind=mesh.POISSON, nx=9, ny=7, nz=5),
    "permuted": dict(kind=mesh.POISSON, nx=10, ny=8, nz=6, permute=True),
    "momentum": dict(kind=mesh.MOMENTUM, nx=8, beta=2.0, jitter=0.1, permute=True,
                     pin_first=False, rhs_zero_mean=False),
    "large": dict(kind=mesh.POISSON, nx=96, permute=True),
}


@pytest.mark.parametrize("name", list(SYNTH))
@pytest.mark.parametrize("base", [0, 1])
def test_synthetic(torch, S, name, base):
    c = mesh.make_case(with_faces=True, **SYNTH[name])
    print(f"topology {name} n={c.n} faces={c.faces.size // 2} base={base}", flush=True)
    check(torch, S, c.n, c.faces, base)


@pytest.mark.parametrize("name", ["cavity", "backstep_nonorthogonal"])
def test_reference_meshes(torch, S, name, tmp_path):
    with gzip.open(GOLDEN / (name + ".cns.gz"), "rb") as src, \
            open(tmp_path / (name + ".cns"), "wb") as dst:
        shutil.copyfileobj(src, dst)
    cns = mesh.read_cns(tmp_path / (name + ".cns"))
    print(f"topology {name} n={cns['n_cells']}", flush=True)
    check(torch, S, cns["n_cells"], cns["faces_c"], 1)


def test_edge_cases(torch, S):
    print("topology edge cases", flush=True)
    check(torch, S, 5, np.zeros(0, np.int32), 1)                      # no faces
    check(torch, S, 3, [1, -1, 2, -2, 3, -3, 3, -4], 0)               # boundary only
    check(torch, S, 4, [1, 2, 2, 3, 1, 2, 3, 4, 2, 1, 4, -1], 1)      # repeated pairs
    check(torch, S, 3, [1, 2, 2, 2, 2, 3], 1)                         # a face of 2 with itself
    rng = np.random.default_rng(4)
    n = 3000
    c1 = rng.integers(1, n + 1, 20000)
    c2 = rng.integers(-50, n + 1, 20000)
    c2[c2 == 0] = -1
    check(torch, S, n, np.stack([c1, c2], 1).ravel(), 1)               # random incidence


def test_errors(torch, S):
    print("topology errors", flush=True)
    check(torch, S, 3, [1, 0, 1, 2], 1)                                # c2 = 0: a boundary
    for bad in ([1, 4], [0, 2], [5, -1]):                              # n_cells = 3
        with pytest.raises(S.SolverError):
            S.topology_device(3, torch.tensor(bad, dtype=torch.int32, device="cuda"))
    lib = S._lib()
    faces = torch.tensor([1, 2, 2, 3], dtype=torch.int32, device="cuda")
    out = [torch.empty(8, dtype=torch.int32, device="cuda") for _ in range(4)]
    nnz = ctypes.c_longlong(0)
    rc = lib.tfs_topology_build_device(3, 2, faces.data_ptr(), 1, 6, out[0].data_ptr(),
                                       out[1].data_ptr(), out[2].data_ptr(), out[3].data_ptr(),
                                       ctypes.byref(nnz))
    assert rc == S.TFS_ERR_INVALID and nnz.value == 7                  # needs 3 + 2 x 2
    big = torch.zeros(6, dtype=torch.int32, device="cuda")
    big[1:5] = faces
    rc = lib.tfs_topology_build_device(3, 2, big[1:].data_ptr(), 1, 7, out[0].data_ptr(),
                                       out[1].data_ptr(), out[2].data_ptr(), None,
                                       ctypes.byref(nnz))
    assert rc == S.TFS_ERR_INVALID                                     # misaligned pairs
