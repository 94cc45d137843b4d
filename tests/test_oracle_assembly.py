"""The face-loop assembly of A % val (oracle/tfs_oracle.c: or_assemble), pinned
on the CPU:
  * bitwise against the reference's own Pressure_Matrix_Fractional (compiled
    from its sources into oracle/_ref) on stretched, jittered meshes with
    real-valued coefficients -- this fixes the order of the diagonal sums;
  * exactly (integer-valued coefficients, so every sum is exact) against a
    dense sum of per-face element matrices [[a12, -a12], [-a21, a21]], with
    boundary adds, cell terms, sav and division by urf -- this fixes where
    every coefficient goes and its sign;
  * bitwise against the synthetic-case generator's own assembly.
"""
import numpy as np
import pytest

from tflows_amd import mesh


def _case(**kw):
    return mesh.make_case(with_faces=True, **kw)


def _topo(c):
    row, col, dia, pos = mesh.topology(c.n, c.faces, base=1, with_pos=True)
    return row, col, dia, pos


@pytest.mark.parametrize("kw", [dict(kind=mesh.POISSON, nx=6, ny=5, nz=4, beta=2.0, jitter=0.1),
                                dict(kind=mesh.POISSON, nx=5, beta=1.0, permute=True)])
def test_pressure_matrix_matches_reference(reference, oracle, kw):
    c = _case(pin_first=False, **kw)
    rng = np.random.default_rng(3)
    nf = c.faces.size // 2
    scoef = rng.uniform(0.1, 2.0, nf)
    dt = 0.0137
    ref_val = reference.pressure_matrix(c.n, c.faces, dt, scoef)
    row, col, dia, pos = _topo(c)
    bnd_on = np.zeros(nf, np.uint8)          # walls: no diagonal add (sequential build)
    val = oracle.assemble(c.n, c.faces, pos, dia, len(col), dt * scoef, bnd_on=bnd_on)
    assert np.array_equal(val, ref_val)


def test_assembly_exact_against_element_matrices(oracle):
    c = _case(kind=mesh.MOMENTUM, nx=5, ny=4, nz=3, beta=1.5, jitter=0.1, permute=True,
              pin_first=False, rhs_zero_mean=False)
    row, col, dia, pos = _topo(c)
    nf = c.faces.size // 2
    rng = np.random.default_rng(11)
    a12 = rng.integers(1, 50, nf).astype(np.float64)
    a21 = rng.integers(1, 50, nf).astype(np.float64)
    bnd_on = rng.integers(0, 2, nf).astype(np.uint8)
    cell = rng.integers(0, 20, c.n).astype(np.float64)
    urf = 0.5
    val, sav = oracle.assemble(c.n, c.faces, pos, dia, len(col), a12, a21, bnd_on, cell, urf,
                               with_sav=True)
    dense = np.zeros((c.n, c.n))
    for s in range(nf):
        c1, c2 = c.faces[2 * s] - 1, c.faces[2 * s + 1] - 1
        if c2 >= 0:
            dense[c1, c1] += a12[s]
            dense[c1, c2] -= a12[s]
            dense[c2, c2] += a21[s]
            dense[c2, c1] -= a21[s]
        elif bnd_on[s]:
            dense[c1, c1] += a12[s]
    dense[np.arange(c.n), np.arange(c.n)] += cell
    assert np.array_equal(sav, np.diag(dense))
    dense[np.arange(c.n), np.arange(c.n)] /= urf
    got = np.zeros((c.n, c.n))
    for i in range(c.n):
        for k in range(row[i] - 1, row[i + 1] - 1):
            got[i, col[k] - 1] = val[k]
    assert np.array_equal(got, dense)


@pytest.mark.parametrize("kw", [dict(kind=mesh.POISSON, nx=7, ny=6, nz=5, beta=2.0, jitter=0.1),
                                dict(kind=mesh.MOMENTUM, nx=6, beta=2.0, jitter=0.1, permute=True,
                                     rhs_zero_mean=False)])
def test_assembly_matches_case_generator(oracle, kw):
    c = _case(pin_first=False, **kw)
    row, col, dia, pos = _topo(c)
    a12, a21, dextra = c.face_coefs
    val = oracle.assemble(c.n, c.faces, pos, dia, len(col), a12, a21, None, dextra)
    assert np.array_equal(row - 1, c.row) and np.array_equal(col - 1, c.col)
    assert np.array_equal(val, c.val)
