"""The solver and the device assembly on real T-FlowS meshes (the reference
generator's cavity and non-orthogonal backstep, read by the product's
.cns/.geo readers), against the oracle and the reference's own outputs.

Bar as for the synthetic systems: SpMV and preconditioner sweeps bitwise;
solves within +-5% iterations and 1e-8 solution rel-L2 at tol 1e-10 ||b||;
the device-assembled pressure matrix bitwise equal to the host's and to the
reference's Pressure_Matrix_Fractional (through its SHA-256).
"""
import gzip
import hashlib
import json
import pathlib
import shutil

import numpy as np
import pytest

from conftest import rel_l2
from tflows_amd import mesh

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "mesh"
META = json.loads((GOLDEN / "mesh_reference.json").read_text())
DT = META["dt"]
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
def cases_(tmp_path_factory):
    d = tmp_path_factory.mktemp("meshes")
    out = {}
    for name in NAMES:
        for ext in (".cns", ".geo"):
            with gzip.open(GOLDEN / (name + ext + ".gz"), "rb") as src, \
                    open(d / (name + ext), "wb") as dst:
                shutil.copyfileobj(src, dst)
        out[name] = mesh.mesh_case(d / name, dt=DT)
    return out


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("sweep", ["default", "levels"])
def test_spmv_and_sweeps_bitwise(torch, S, oracle, cases_, name, prec, sweep):
    c = cases_[name]
    print(f"mesh {name} n={c.n} prec={prec} sweep={sweep}", flush=True)
    A = S.Matrix(c.row, c.col, c.dia, c.val)
    if sweep == "levels":
        A.set_sweep_schedule(S.TFS_SWEEP_LEVELS)
    x = np.cos(np.arange(c.n) * 0.11)
    y = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    A.spmv(dev(torch, x), y)
    assert np.array_equal(host(y), oracle.spmv(c.row, c.col, c.dia, c.val, x))
    A.prec_form(prec)
    d = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    A.prec_diag(d)
    d_or = oracle.prec_form(c.row, c.col, c.dia, c.val, prec)
    assert np.array_equal(host(d), d_or)
    z = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    A.prec_solve(prec, dev(torch, c.b), z)
    assert np.array_equal(host(z), oracle.prec_solve(c.row, c.col, c.dia, c.val, d_or, prec, c.b))


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("method", ["cg", "bicg", "cgs"])
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_solve_parity(torch, S, oracle, cases_, name, method, prec):
    from test_gpu_parity import iters_close, iters_close_cgs
    c = cases_[name]
    print(f"mesh {name} {method} prec={prec}", flush=True)
    tol = 1e-10 * float(np.linalg.norm(c.b))
    cap = 20000
    ref = oracle.solve(method, c.row, c.col, c.dia, c.val, c.b, prec=prec, niter=cap, tol=tol)
    A = S.Matrix(c.row, c.col, c.dia, c.val)
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    r = dev(torch, c.b)
    niter, ini, fin = S.solve(method, A, x, r, prec, cap, tol)
    if method == "cgs":
        assert iters_close_cgs(oracle, c, prec, tol, niter, ref["niter"]), (niter, ref["niter"])
    else:
        assert iters_close(niter, ref["niter"]), (niter, ref["niter"])
    assert abs(ini - ref["ini_res"]) <= 1e-12 * ref["ini_res"]
    if ref["niter"] <= cap:
        assert fin < tol and A.status == S.TFS_OK
        assert rel_l2(host(x), ref["x"]) <= 1e-8
    else:
        assert niter == cap + 1 and A.status == S.TFS_NOT_CONVERGED


@pytest.mark.parametrize("name", NAMES)
def test_against_reference_solution(torch, S, cases_, name):
    """The reference's own Cg (prec 2) on the same pressure system."""
    c = cases_[name]
    z = np.load(GOLDEN / f"{name}_cg_p2.npz", allow_pickle=False)
    A = S.Matrix(c.row, c.col, c.dia, c.val)
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    r = dev(torch, c.b)
    niter, ini, fin = S.Cg(A, x, r, 2, 20000, float(z["tol"]))
    want = int(z["niter"])
    assert abs(niter - want) <= max(1, want // 20), (niter, want)
    assert abs(ini - float(z["ini_res"])) <= 1e-12 * float(z["ini_res"])
    assert fin < float(z["tol"])
    assert rel_l2(host(x), z["x"]) <= 1e-8


@pytest.mark.parametrize("name", NAMES)
def test_device_assembly_on_mesh(torch, S, cases_, name):
    """Pressure_Matrix_Fractional's face loop on the device from dt*Scoef."""
    c = cases_[name]
    row, col, dia, pos = mesh.topology(c.n, c.faces, base=1, with_pos=True)
    A = S.Matrix(row - 1, col - 1, dia - 1)
    asm = S.Assembly(A, c.faces, pos, dia, base=1)
    a12 = c.face_coefs[0]
    asm.assemble(dev(torch, a12))
    got = A.get_values()
    want = c.val.copy()
    want[c.dia[0]] /= 2.0            # mesh_case pins cell 1 after assembling
    assert np.array_equal(got, want)
    h = hashlib.sha256(np.ascontiguousarray(got).tobytes()).hexdigest()
    assert h == META["meshes"][name]["sha256"]["val"]
