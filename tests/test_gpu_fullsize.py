"""Configurations 3 and 4 at their full sizes, through size-independent
properties (the oracle would take minutes to hours here):

* Poisson 256^3 stretched to cell aspect ratio 100 (16,777,216 rows): the
  FP32-inner refinement (reliable updates, the default) returns a true FP64
  residual below tol (checked here with the library's own FP64 residual), and
  its solution agrees with the plain FP64 Cg drop-in's on the same system to
  rel-L2 <= 1e-5 (the FP32 recurrence cannot see residual components below FP32
  resolution, so at equal residual the two differ along the smallest
  eigenvectors: 2e-7 - 1.2e-6 in the measured runs), the FP64 Cg reaching its
  tol too.
* The 64 uneven Poisson systems of the batch configuration (5.4e6 rows in
  all): every system's Cg and FP32-inner refinement end with status OK and a
  final residual below its own tol.
"""
import numpy as np
import pytest

from tflows_amd import cases, mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.fail("no GPU visible")
    return t


def test_poisson256_refinement_full_size(torch):
    from tflows_amd import solvers as S
    kw = cases.workload_kwargs(3)
    c = mesh.make_case(**kw)
    print(f"Poisson {kw['nx']}^3 n={c.n} nnz={c.nnz}", flush=True)
    tol = 1e-10 * float(np.linalg.norm(c.b))
    A = S.Matrix(c.row, c.col, c.dia, c.val)
    b = torch.from_numpy(c.b).cuda()
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    r = b.clone()
    steps, inner, ini, fin = S.solve_mixed(A, x, r, 2, 200, tol)
    print(f"  refinement: {steps} replacements, {inner} FP32 iterations, fin {fin / tol:.3f} tol",
          flush=True)
    assert A.status == S.TFS_OK and fin < tol
    # the default inner order at this size: 16^3 bricks in two colours
    info = A.info()
    assert (info["inner_order"], info["inner_brick_edge"], info["inner_colours"]) == (
        S.TFS_INNER_BRICKS, 16, 2), info
    tr = b.clone()
    A.residual(x, tr)  # the true FP64 residual, independently of the solve
    assert float(torch.linalg.norm(tr)) < tol
    assert abs(float(torch.linalg.norm(tr)) - fin) <= 1e-3 * tol
    x32 = x.cpu().numpy()
    x.zero_()
    r.copy_(b)
    niter, _, fin64 = S.Cg(A, x, r, 2, 100000, tol)
    assert A.status == S.TFS_OK and fin64 < tol
    x64 = x.cpu().numpy()
    err = float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64))
    print(f"  FP64 Cg {niter} iterations; rel-L2 refinement vs FP64 {err:.2e}", flush=True)
    assert err <= 1e-5


def test_batch64_full_size(torch):
    from tflows_amd import solvers as S
    sysl = [mesh.make_case(**kw) for kw in cases.batch_systems()]
    B = S.Batch([(c.row, c.col, c.dia, c.val) for c in sysl])
    print("batch", B.info(), flush=True)
    tols = np.array([1e-10 * float(np.linalg.norm(c.b)) for c in sysl])

    def fresh():
        return ([torch.zeros(c.n, dtype=torch.float64, device="cuda") for c in sysl],
                [torch.from_numpy(c.b.copy()).cuda() for c in sysl])
    xs, rs = fresh()
    niter, ini, fin, st = B.solve(xs, rs, 2, [20000] * len(sysl), tols)
    assert np.all(st == S.TFS_OK) and np.all(fin < tols), (st, fin / tols)
    xs, rs = fresh()
    steps, inner, ini, fin, st = B.solve_mixed(xs, rs, 2, 200, tols)
    print(f"  refinement: steps {int(steps.max())} max, inner {int(inner.sum())} in all", flush=True)
    assert np.all(st == S.TFS_OK) and np.all(fin < tols), (st, fin / tols)
