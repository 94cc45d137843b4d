"""FP32-inner PCG inside FP64 iterative refinement against the FP64 oracle.

The refinement solver has no counterpart in the reference; what pins it is the
FP64 answer.  The refined solution must reach the same FP64 tolerance (the
final residual is the true ||b - A x||, computed in FP64).  Its error against
the exact solution (sparse LU) is printed for every case and bounded twice:
  * stopping on the residual alone (the default), within 1e-6: each step
    contracts the error by about cond(A) * 2^-24, so on the aspect-100 Poisson
    case the residual passes 1e-10 ||b|| while the error is ~5e-8, where the
    oracle's FP64 Cg iterate is at ~1e-10;
  * with the flexible-CG outer loop (the default) and the update criterion
    x_rtol = 1e-10, as accurate as the oracle's FP64 Cg at the same
    tolerance: within 10x of the oracle's own error and <= 1e-8 wherever the
    oracle's is (defect correction's update does not contain the error it
    cannot resolve, so there the bound stays that of the residual).
The reliable-update loop (the default) runs one FP32 CG recurrence with FP64
residual replacement: in the reference order it takes about the oracle's
iteration count; its error meets the residual bound above, and with the
correction criterion (the last segment's FP64 correction <= 1e-10 ||x||) the
oracle's own accuracy.
The FP32 factor is the reference recurrence on FP32 values (FP64 arithmetic,
FP32 storage): it agrees with the FP64 oracle factor to FP32 rounding (rel
1e-5).
"""
import numpy as np
import pytest

from conftest import exact_solution, refinement_bar, rel_l2
from tflows_amd import cases, mesh

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
    "poisson_stretched20": dict(kind=mesh.POISSON, nx=20, beta=cases.beta_for_aspect(20, 100.0)),
    "momentum12": dict(kind=mesh.MOMENTUM, nx=12, beta=2.0, jitter=0.1, permute=True,
                       pin_first=False, rhs_zero_mean=False),
    "poisson_perm": dict(kind=mesh.POISSON, nx=14, ny=10, nz=8, beta=1.5, permute=True),
}


OUTER_CRITERIA = [("reliable", "residual"), ("reliable", "correction"), ("fcg", "residual"),
                  ("fcg", "correction"), ("defect", "residual"), ("defect", "correction")]


ORDERS = {"reference": "TFS_INNER_REFERENCE", "coloured": "TFS_INNER_COLOURED",
          "bricks": "TFS_INNER_BRICKS"}


@pytest.mark.parametrize("outer,criterion", OUTER_CRITERIA)
@pytest.mark.parametrize("order", list(ORDERS))
@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_refinement_reaches_fp64_solution(torch, S, oracle, name, prec, order, criterion, outer):
    c = mesh.make_case(**CASES[name])
    print(f"case {name} prec={prec} order={order} criterion={criterion} outer={outer} n={c.n}",
          flush=True)
    tol = 1e-10 * float(np.linalg.norm(c.b))
    ref = oracle.solve("cg", c.row, c.col, c.dia, c.val, c.b, prec=prec, niter=20000, tol=tol)
    A = S.Matrix(c.row, c.col, c.dia, c.val)
    A.set_inner_order(getattr(S, ORDERS[order]))
    A.set_refinement({"reliable": S.TFS_REFINE_RELIABLE, "fcg": S.TFS_REFINE_FCG,
                      "defect": S.TFS_REFINE_DEFECT}[outer])
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    r = torch.from_numpy(c.b.copy()).cuda()
    x_rtol = 1e-10 if criterion == "correction" else 0.0
    rtol = 0.1 if outer == "reliable" else 1e-4
    steps, inner, ini, fin = S.solve_mixed(A, x, r, prec, 50, tol, inner_rtol=rtol, x_rtol=x_rtol)
    if outer == "reliable" and ref["niter"] > 20000:
        # CG itself does not converge here (unpreconditioned nonsymmetric
        # momentum matrix): neither does its FP32 recurrence, and the solve
        # says so within its segment cap
        print(f"  oracle Cg did not converge; reliable: steps {steps} inner {inner}", flush=True)
        assert A.status == S.TFS_NOT_CONVERGED and steps == 51 and inner <= 20000 * 50
        return
    assert A.status == S.TFS_OK and fin < tol, (steps, inner, fin, tol)
    assert 1 <= steps <= 30
    assert abs(ini - np.linalg.norm(c.b)) <= 1e-12 * np.linalg.norm(c.b)
    xs = exact_solution(c.row, c.col, c.val, c.b)
    err, err_ref = rel_l2(x.cpu().numpy(), xs), rel_l2(ref["x"], xs)
    print(f"  steps {steps} inner {inner} (oracle {ref['niter']}): rel-L2 vs exact {err:.2e} "
          f"(oracle {err_ref:.2e}), vs oracle {rel_l2(x.cpu().numpy(), ref['x']):.2e}", flush=True)
    if outer == "reliable" and order == "reference" and criterion == "residual":
        # one CG recurrence whose preconditioner is the oracle's in FP32:
        # about the oracle's iteration count
        assert inner <= 1.5 * ref["niter"] + 5, (inner, ref["niter"])
    if criterion == "correction" and outer in ("fcg", "reliable"):
        assert err <= refinement_bar(ref["x"], xs)
    else:
        # the residual alone does not fix the error along the smallest
        # eigenvectors, which FP32 resolves less well than FP64 at the same
        # residual; and defect correction's update misses what FP32 cannot
        # resolve, so there the update criterion does not bound the error
        # either: the stated bound is that of the residual
        assert err <= max(1e-6, refinement_bar(ref["x"], xs))
    # the residual left in r1 is the true FP64 residual
    true_r = torch.from_numpy(c.b.copy()).cuda()
    A.residual(x, true_r)
    assert float(torch.linalg.norm(true_r - r)) <= 1e-12 * float(np.linalg.norm(c.b))


@pytest.mark.parametrize("name", ["poisson_stretched20", "momentum12"])
def test_fp32_factor_is_the_reference_recurrence(torch, S, oracle, name):
    c = mesh.make_case(**CASES[name])
    A = S.Matrix(c.row, c.col, c.dia, c.val)
    d32 = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    A.prec_diag32(2, d32)
    d64 = oracle.prec_form(c.row, c.col, c.dia, c.val, 2)
    np.testing.assert_allclose(d32.cpu().numpy(), d64, rtol=1e-5)
    # and it is really FP32: every entry is representable in float
    v = d32.cpu().numpy()
    assert np.array_equal(v.astype(np.float32).astype(np.float64), v)


def test_refinement_cap_and_early_exit(torch, S):
    c = mesh.make_case(**CASES["poisson_perm"])
    A = S.Matrix(c.row, c.col, c.dia, c.val)
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    r = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    steps, inner, ini, fin = S.solve_mixed(A, x, r, 2, 10, 1e-8)
    assert steps == 0 and inner == 0 and fin == 0.0
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    r = torch.from_numpy(c.b.copy()).cuda()
    steps, inner, ini, fin = S.solve_mixed(A, x, r, 2, 1, 1e-14 * float(np.linalg.norm(c.b)))
    assert steps == 2 and A.status == S.TFS_NOT_CONVERGED


def test_coloured_inner_order(torch, S):
    """A hex mesh in lexicographic order is 2-colourable: the inner sweeps have
    two levels; the refined answer does not depend on the inner order beyond
    the FP64 tolerance, and a new set of values reaches the inner copy."""
    c = mesh.make_case(**CASES["poisson_stretched20"])
    tol = 1e-10 * float(np.linalg.norm(c.b))
    A = S.Matrix(c.row, c.col, c.dia, c.val)
    A.set_inner_block(1)  # point colouring
    out = {}
    for order in (S.TFS_INNER_REFERENCE, S.TFS_INNER_COLOURED):
        A.set_inner_order(order)
        x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
        r = torch.from_numpy(c.b.copy()).cuda()
        steps, inner, ini, fin = S.solve_mixed(A, x, r, 2, 50, tol, x_rtol=1e-10)
        assert A.status == S.TFS_OK and fin < tol
        out[order] = x.cpu().numpy()
        print("order", order, "steps", steps, "inner", inner, flush=True)
    assert A.info()["inner_colours"] == 2
    xs = exact_solution(c.row, c.col, c.val, c.b)
    ref = S.Matrix(c.row, c.col, c.dia, c.val)
    x64 = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    r64 = torch.from_numpy(c.b.copy()).cuda()
    S.Cg(ref, x64, r64, 2, 20000, tol)
    bar = refinement_bar(x64.cpu().numpy(), xs)
    for order in out:
        assert rel_l2(out[order], xs) <= bar
    # new values reach the colour-ordered copy: doubled values, half the solution
    A.set_values(2.0 * c.val)
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    r = torch.from_numpy(c.b.copy()).cuda()
    steps, inner, ini, fin = S.solve_mixed(A, x, r, 2, 50, tol / 2, x_rtol=1e-10)
    assert A.status == S.TFS_OK
    assert rel_l2(2.0 * x.cpu().numpy(), xs) <= bar


@pytest.mark.parametrize("block", [4, 16])
def test_block_coloured_inner_order(torch, S, oracle, block):
    """Block multi-coloured inner order: the refined answer meets the same bar
    against the exact solution, and changing the block size rebuilds the copy."""
    c = mesh.make_case(**CASES["poisson_stretched20"])
    tol = 1e-10 * float(np.linalg.norm(c.b))
    ref = oracle.solve("cg", c.row, c.col, c.dia, c.val, c.b, prec=2, niter=20000, tol=tol)
    xs = exact_solution(c.row, c.col, c.val, c.b)
    A = S.Matrix(c.row, c.col, c.dia, c.val)
    inners = []
    for blk in (1, block):
        A.set_inner_block(blk)
        x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
        r = torch.from_numpy(c.b.copy()).cuda()
        steps, inner, ini, fin = S.solve_mixed(A, x, r, 2, 50, tol, x_rtol=1e-10)
        assert A.status == S.TFS_OK and fin < tol
        assert rel_l2(x.cpu().numpy(), xs) <= refinement_bar(ref["x"], xs)
        inners.append(inner)
        print("block", blk, "colours", A.info()["inner_colours"], "steps", steps, "inner", inner,
              flush=True)
    assert A.info()["inner_colours"] >= 2


def test_brick_inner_order(torch, S):
    """The brick order on a lexicographically numbered box: two colours of
    4^3 bricks, the refined answer meets the same bar against the exact
    solution, and the stronger block factor takes fewer FP32 iterations than
    point colouring (reliable updates, prec 2)."""
    c = mesh.make_case(nx=24, kind=mesh.POISSON, beta=cases.beta_for_aspect(24, 100.0))
    tol = 1e-10 * float(np.linalg.norm(c.b))
    xs = exact_solution(c.row, c.col, c.val, c.b)
    inner = {}
    for order in ("coloured", "bricks"):
        A = S.Matrix(c.row, c.col, c.dia, c.val)
        A.set_inner_order(getattr(S, ORDERS[order]))
        x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
        r = torch.from_numpy(c.b.copy()).cuda()
        steps, inner[order], ini, fin = S.solve_mixed(A, x, r, 2, 50, tol)
        assert A.status == S.TFS_OK and fin < tol
        assert A.info()["inner_colours"] == 2
        assert A.info()["inner_brick_edge"] == (4 if order == "bricks" else 0)
        assert rel_l2(x.cpu().numpy(), xs) <= 1e-6
        print(order, "steps", steps, "inner", inner[order], "levels", A.info()["inner_levels"],
              flush=True)
    assert inner["bricks"] < 0.8 * inner["coloured"], inner
