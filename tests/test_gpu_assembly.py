"""Device-side assembly of A % val (tfs_assembly_*) against the oracle's
face-loop restatement (itself pinned bitwise on the CPU against the
reference's Pressure_Matrix_Fractional): bitwise for real-valued
coefficients, with and without a21, boundary adds, cell terms, sav and
under-relaxation; and a solve after the device assembly equals, bitwise, the
solve after uploading the same values from the host (and meets the parity
bar against the oracle's solve)."""
import numpy as np
import pytest

from conftest import rel_l2
from tflows_amd import mesh

pytestmark = pytest.mark.gpu


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


CASES = {
    "momentum_perm": dict(kind=mesh.MOMENTUM, nx=9, ny=7, nz=6, beta=2.0, jitter=0.1,
                          permute=True, pin_first=False, rhs_zero_mean=False),
    "poisson_box": dict(kind=mesh.POISSON, nx=12, beta=1.5, pin_first=False),
}


def _setup(c):
    row, col, dia, pos = mesh.topology(c.n, c.faces, base=1, with_pos=True)
    return row, col, dia, pos


@pytest.mark.parametrize("variant", ["pressure", "momentum", "relaxed"])
@pytest.mark.parametrize("name", list(CASES))
def test_device_assembly_bitwise(torch, S, oracle, name, variant):
    c = mesh.make_case(with_faces=True, **CASES[name])
    row, col, dia, pos = _setup(c)
    nf = c.faces.size // 2
    rng = np.random.default_rng(7)
    a12 = rng.uniform(0.01, 3.0, nf)
    a21 = None if variant == "pressure" else rng.uniform(0.01, 3.0, nf)
    bnd = None if variant == "pressure" else rng.integers(0, 2, nf).astype(np.uint8)
    cell = None if variant == "pressure" else rng.uniform(0.0, 5.0, c.n)
    urf = 0.7 if variant == "relaxed" else 1.0
    print(f"case {name} variant {variant} n={c.n} faces={nf}", flush=True)
    want, want_sav = oracle.assemble(c.n, c.faces, pos, dia, len(col), a12, a21, bnd, cell, urf,
                                     with_sav=True)
    A = S.Matrix(row - 1, col - 1, dia - 1)
    asm = S.Assembly(A, c.faces, pos, dia, base=1)

    def dev(a, dt=torch.float64):
        return None if a is None else torch.from_numpy(a).to("cuda", dt)
    sav = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    asm.assemble(dev(a12), dev(a21), dev(bnd, torch.uint8), dev(cell), urf, sav)
    assert np.array_equal(A.get_values(), want)
    assert np.array_equal(sav.cpu().numpy(), want_sav)


def test_solve_after_device_assembly(torch, S, oracle):
    c = mesh.make_case(with_faces=True, **CASES["momentum_perm"])
    row, col, dia, pos = _setup(c)
    a12, a21, dextra = c.face_coefs
    A = S.Matrix(row - 1, col - 1, dia - 1)
    asm = S.Assembly(A, c.faces, pos, dia, base=1)
    asm.assemble(torch.from_numpy(a12).cuda(), torch.from_numpy(a21).cuda(), None,
                 torch.from_numpy(dextra).cuda())
    B = S.Matrix(c.row, c.col, c.dia, c.val)    # the same values from the host
    tol = 1e-10 * float(np.linalg.norm(c.b))
    out = []
    for M in (A, B):
        x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
        r = torch.from_numpy(c.b.copy()).cuda()
        it, ini, fin = S.Bicg(M, x, r, 2, 5000, tol)
        out.append((it, fin, x.cpu().numpy()))
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])
    ref = oracle.solve("bicg", c.row, c.col, c.dia, c.val, c.b, prec=2, niter=5000, tol=tol)
    assert abs(out[0][0] - ref["niter"]) <= max(1, ref["niter"] // 20)
    assert rel_l2(out[0][2], ref["x"]) <= 1e-8
