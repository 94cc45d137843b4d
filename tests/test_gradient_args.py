"""The mesh-side bindings check every per-face and per-cell array against the
mesh before any device access (CPU only): a flux shorter or longer than the
mesh's face count, or surface vectors not exactly 3 x n_faces, would make the
face kernels read or write past the arrays (or with the wrong stride)."""
import numpy as np
import pytest

from tflows_amd import solvers


def mesh_object(n=10, nb=4, nf=17):
    g = solvers.Gradient.__new__(solvers.Gradient)  # no device: sizes only
    g.n, g.nb, g.nf = n, nb, nf
    return g


def pf_args(g, nf_flux=None, s_len=None):
    n, na, nf = g.n, g.n + g.nb, g.nf
    z = np.zeros
    return dict(f=z(nf), s=z(3 * nf if s_len is None else s_len), vol=z(n), sav=z(n), u=z(na),
                v=z(na), w=z(na), px=z(n), py=z(n), pz=z(n), p=z(na), den=z(na),
                flux=z(nf if nf_flux is None else nf_flux), a12=z(nf), b=z(n))


@pytest.mark.parametrize("nf_flux", [16, 18])
def test_pressure_fractional_flux_length(nf_flux):
    g = mesh_object()
    with pytest.raises(ValueError, match="flux"):
        g.pressure_fractional(**pf_args(g, nf_flux=nf_flux))


@pytest.mark.parametrize("s_len", [3 * 17 - 1, 3 * 17 + 3])
def test_pressure_fractional_surface_vectors(s_len):
    g = mesh_object()
    with pytest.raises(ValueError, match="sx, sy, sz"):
        g.pressure_fractional(**pf_args(g, s_len=s_len))


@pytest.mark.parametrize("nf_flux", [16, 18])
def test_correct_velocity_flux_length(nf_flux):
    g = mesh_object()
    n, na, nf = g.n, g.n + g.nb, g.nf
    z = np.zeros
    with pytest.raises(ValueError, match="flux"):
        g.correct_velocity(vol=z(n), sav=z(n), px=z(n), py=z(n), pz=z(n), u=z(na), v=z(na),
                           w=z(na), pp=z(na), coef=z(nf), den=z(n), scoef=z(nf), dt=1.0,
                           visc=1.0, flux=z(nf_flux), b=z(n))


def test_short_cell_array():
    g = mesh_object()
    args = pf_args(g)
    args["u"] = np.zeros(g.n)  # cells only: boundary cells missing
    with pytest.raises(ValueError, match="u:"):
        g.pressure_fractional(**args)
