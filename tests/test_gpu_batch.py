"""Batched solves (many independent systems in one block-diagonal matrix).

Each system of the batch must give what it gives in a batch of its own: the
same iteration count and, bitwise, the same solution (every kernel's
arithmetic per row and the per-system reductions, in tile order, are
independent of the other systems).  Against the single-matrix solve (whose
dot products are summed in another order) and the oracle it meets the parity
bar of the single solve.
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


def small_batch(k=6):
    out = []
    for s, kw in enumerate(cases.batch_systems(count=k, lo=5, hi=14)):
        out.append(mesh.make_case(**kw))
    out.append(mesh.make_case(kind=mesh.MOMENTUM, nx=7, beta=2.0, jitter=0.1, permute=True,
                              pin_first=False, rhs_zero_mean=False, seed=99))
    return out


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_batch_equals_single_solves(torch, S, oracle, prec):
    sysl = small_batch()
    B = S.Batch([(c.row, c.col, c.dia, c.val) for c in sysl])
    tols = [1e-10 * float(np.linalg.norm(c.b)) for c in sysl]
    xs = [torch.zeros(c.n, dtype=torch.float64, device="cuda") for c in sysl]
    rs = [torch.from_numpy(c.b.copy()).cuda() for c in sysl]
    niter, ini, fin, st = B.solve(xs, rs, prec, [3000] * len(sysl), tols)
    print("batch niter", niter.tolist(), "info", B.info(), flush=True)
    for k, c in enumerate(sysl):
        B1 = S.Batch([(c.row, c.col, c.dia, c.val)])
        xb = [torch.zeros(c.n, dtype=torch.float64, device="cuda")]
        rb = [torch.from_numpy(c.b.copy()).cuda()]
        nb, inib, finb, stb = B1.solve(xb, rb, prec, [3000], [tols[k]])
        assert niter[k] == nb[0] and st[k] == stb[0]
        assert np.array_equal(xs[k].cpu().numpy(), xb[0].cpu().numpy()), k
        assert fin[k] == finb[0] and ini[k] == inib[0]
        A = S.Matrix(c.row, c.col, c.dia, c.val)
        x1 = torch.zeros(c.n, dtype=torch.float64, device="cuda")
        r1 = torch.from_numpy(c.b.copy()).cuda()
        it1, ini1, fin1 = S.Cg(A, x1, r1, prec, 3000, tols[k])
        ref = oracle.solve("cg", c.row, c.col, c.dia, c.val, c.b, prec=prec, niter=3000,
                           tol=tols[k])
        assert abs(int(niter[k]) - it1) <= max(1, it1 // 20)
        assert abs(niter[k] - ref["niter"]) <= max(1, ref["niter"] // 20)
        if ref["niter"] > 3000:   # the reference does not converge either (Cg, nonsymmetric)
            assert st[k] == S.TFS_NOT_CONVERGED
            continue
        assert st[k] == S.TFS_OK
        assert rel_l2(xs[k].cpu().numpy(), x1.cpu().numpy()) <= 1e-8
        assert rel_l2(xs[k].cpu().numpy(), ref["x"]) <= 1e-8
    assert B.info()["graph_launches"] > 0


REFINE = {"reliable": "TFS_REFINE_RELIABLE", "fcg": "TFS_REFINE_FCG", "defect": "TFS_REFINE_DEFECT"}


@pytest.mark.parametrize("outer", ["reliable", "fcg", "defect"])
@pytest.mark.parametrize("order", ["reference", "coloured", "bricks"])
def test_batch_mixed_refinement(torch, S, oracle, order, outer):
    """Every system reaches its FP64 tolerance (true residual), with the error
    against the exact solution bounded as for the single refinement solve
    (stopping on the residual: within 1e-6, or within 10x of the FP64
    oracle's own error where that is larger).  The colour- or brick-ordered
    inner copy is the same batch of systems, so its results stay per system.
    (With the reference order the reliable-update setting runs flexible CG.)"""
    sysl = small_batch()
    B = S.Batch([(c.row, c.col, c.dia, c.val) for c in sysl])
    orders = {"reference": S.TFS_INNER_REFERENCE, "coloured": S.TFS_INNER_COLOURED,
              "bricks": S.TFS_INNER_BRICKS}
    B.set_inner_order(orders[order])
    B.set_refinement(getattr(S, REFINE[outer]))
    tols = [1e-10 * float(np.linalg.norm(c.b)) for c in sysl]
    xs = [torch.zeros(c.n, dtype=torch.float64, device="cuda") for c in sysl]
    rs = [torch.from_numpy(c.b.copy()).cuda() for c in sysl]
    steps, inner, ini, fin, st = B.solve_mixed(xs, rs, 2, 50, tols)
    print("order", order, "outer", outer, "steps", steps.tolist(), "inner", inner.tolist(),
          flush=True)
    for k, c in enumerate(sysl):
        assert st[k] == S.TFS_OK and fin[k] < tols[k]
        ref = oracle.solve("cg", c.row, c.col, c.dia, c.val, c.b, prec=2, niter=5000, tol=tols[k])
        xe = exact_solution(c.row, c.col, c.val, c.b)
        err = rel_l2(xs[k].cpu().numpy(), xe)
        print(f"  system {k}: rel-L2 vs exact {err:.2e} (oracle {rel_l2(ref['x'], xe):.2e})",
              flush=True)
        assert err <= max(1e-6, refinement_bar(ref["x"], xe))
    if order != "reference":
        # the same systems alone in a batch give the same results, bitwise
        B1 = S.Batch([(sysl[0].row, sysl[0].col, sysl[0].dia, sysl[0].val)])
        B1.set_inner_order(orders[order])
        B1.set_refinement(getattr(S, REFINE[outer]))
        x1 = [torch.zeros(sysl[0].n, dtype=torch.float64, device="cuda")]
        r1 = [torch.from_numpy(sysl[0].b.copy()).cuda()]
        s1, i1, _, f1, st1 = B1.solve_mixed(x1, r1, 2, 50, tols[:1])
        assert s1[0] == steps[0] and i1[0] == inner[0] and f1[0] == fin[0]
        assert np.array_equal(x1[0].cpu().numpy(), xs[0].cpu().numpy())


def test_batch_caps_and_early_exit(torch, S):
    sysl = small_batch(3)
    B = S.Batch([(c.row, c.col, c.dia, c.val) for c in sysl])
    xs = [torch.zeros(c.n, dtype=torch.float64, device="cuda") for c in sysl]
    rs = [torch.from_numpy(c.b.copy()).cuda() for c in sysl]
    rs[1].zero_()  # ||b|| = 0: early exit, x untouched
    xs[1].fill_(2.0)
    caps = [3, 100, 0, 1000]
    tols = [1e-14 * float(np.linalg.norm(c.b)) for c in sysl]
    tols[3] = 1e-10 * float(np.linalg.norm(sysl[3].b))
    niter, ini, fin, st = B.solve(xs, rs, 2, caps, tols)
    assert niter[0] == 4 and st[0] == S.TFS_NOT_CONVERGED
    assert niter[1] == 0 and bool(torch.all(xs[1] == 2.0))
    assert niter[2] == 1 and st[2] == S.TFS_NOT_CONVERGED
    assert st[3] == S.TFS_OK and fin[3] < tols[3]
