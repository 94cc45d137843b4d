"""Case generator and topology builder (host code, no GPU)."""
import numpy as np
import pytest

from tflows_amd import cases, mesh


def test_topology_matches_compiled_reference(reference):
    c = mesh.make_case(kind=mesh.MOMENTUM, nx=9, ny=6, nz=5, beta=2.0, jitter=0.1,
                       permute=True, with_faces=True)
    t = reference.topology(c.n, c.faces)
    row, col, dia, pos = mesh.topology(c.n, c.faces, base=1, with_pos=True)
    assert np.array_equal(row, t["row"]) and np.array_equal(col, t["col"])
    assert np.array_equal(dia, t["dia"]) and np.array_equal(pos, t["pos"])


def test_topology_matches_oracle(oracle):
    c = mesh.make_case(kind=mesh.POISSON, nx=6, ny=4, nz=5, permute=True, with_faces=True)
    t = oracle.topology(c.n, c.faces)
    row, col, dia, pos = mesh.topology(c.n, c.faces, base=1, with_pos=True)
    for a, k in ((row, "row"), (col, "col"), (dia, "dia"), (pos, "pos")):
        assert np.array_equal(a, t[k]), k


def test_topology_rejects_bad_cells():
    with pytest.raises(ValueError):
        mesh.topology(3, np.array([1, 5], np.int32))


@pytest.mark.parametrize("n", [2, 5, 32])
def test_sizes_of_the_7_point_box(n):
    c = mesh.make_case(kind=mesh.POISSON, nx=n)
    assert c.n == n ** 3
    assert c.nnz == n ** 3 + 6 * n * n * (n - 1)


def test_poisson_matrix_properties():
    c = mesh.make_case(kind=mesh.POISSON, nx=6, beta=1.5, permute=True, pin_first=True)
    A = np.zeros((c.n, c.n))
    for i in range(c.n):
        A[i, c.col[c.row[i]:c.row[i + 1]]] = c.val[c.row[i]:c.row[i + 1]]
    assert np.allclose(A, A.T)
    rs = A.sum(axis=1)
    assert abs(rs[0] - c.val[c.dia[0]] / 2) < 1e-12  # pinned: diagonal doubled
    assert np.abs(rs[1:]).max() < 1e-12              # Neumann walls: zero row sums
    assert np.all(np.linalg.eigvalsh(A) > 0)
    assert abs(c.b.sum()) < 1e-12


def test_momentum_matrix_is_an_m_matrix():
    c = mesh.make_case(kind=mesh.MOMENTUM, nx=6, beta=2.0, jitter=0.1, permute=True,
                       pin_first=False, rhs_zero_mean=False)
    off = np.ones(c.nnz, bool)
    off[c.dia] = False
    assert np.all(c.val[off] <= 0)
    diag = c.val[c.dia]
    offsum = np.array([-c.val[c.row[i]:c.row[i + 1]][off[c.row[i]:c.row[i + 1]]].sum()
                       for i in range(c.n)])
    assert np.all(diag > offsum)  # strict diagonal dominance from the V/dt term
    A = np.zeros((c.n, c.n))
    for i in range(c.n):
        A[i, c.col[c.row[i]:c.row[i + 1]]] = c.val[c.row[i]:c.row[i + 1]]
    assert np.abs(A - A.T).max() > 1e-6
    assert c.dt > 0


def test_cases_are_deterministic():
    a = mesh.make_case(**cases.workload_kwargs(2, scale=0.1))
    b = mesh.make_case(**cases.workload_kwargs(2, scale=0.1))
    assert np.array_equal(a.val, b.val) and np.array_equal(a.col, b.col)
    assert np.array_equal(a.b, b.b)


def test_aspect_ratio_of_the_stretched_poisson():
    n = 32
    beta = cases.beta_for_aspect(n, 100.0)
    c = mesh.make_case(kind=mesh.POISSON, nx=n, beta=beta)
    assert 90 < c.max_aspect < 110


def test_batch_sizes_in_range():
    sys_ = cases.batch_systems()
    assert len(sys_) == 64
    for s in sys_:
        for k in ("nx", "ny", "nz"):
            assert 16 <= s[k] <= 96
        assert 1.0 <= s["beta"] <= 3.0


def test_rng_is_uniform():
    u = np.array([mesh.rng_u01(7, 1, i) for i in range(20000)])
    assert 0 <= u.min() and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01
