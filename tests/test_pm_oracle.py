"""The PM oracle (oracle/pm.py), pinned by what the mathematics fixes: the
long-range field of a point mass (the complement of the short-range split factor),
exact momentum conservation of CIC with an antisymmetric gradient, a vanishing
force on a uniform lattice, linearity in mass and translation by a mesh cell.
The reference's PM (src/partmesh.c, src/conv.f90) needs MPI and 2DECOMP&FFT and
cannot be built here; these checks stand in for it."""
import numpy as np
import pytest

import pm

N, L = 32, 100000.0
H = L / N
RS = 1.25 * H


def field(src, mass, probes):
    phi = pm.potential(pm.deposit(src, mass, N, L), L, RS)
    return pm.interpolate(pm.gradient(phi, N, L), probes, N, L)


def test_point_mass_long_range_field():
    src = np.array([[0.37 * L, 0.41 * L, 0.53 * L]])
    d = np.array([0.6, 0.7, 0.39])
    d /= np.linalg.norm(d)
    r = np.array([2.0, 3.0, 4.0, 6.0]) * H
    a = field(src, 1.0, src + r[:, None] * d)
    radial = -(a @ d)  # attraction points back to the source
    want = pm.long_range_point(r, 1.0, RS)
    assert np.all(np.abs(radial / want - 1) < 0.04), radial / want
    # transverse part small
    trans = np.linalg.norm(a + radial[:, None] * d, axis=1)
    assert np.all(trans < 0.05 * np.abs(radial))


def test_momentum_is_conserved():
    rng = np.random.Generator(np.random.PCG64(1))
    pos = rng.random((400, 3)) * L
    acc = pm.pm_force(pos, 2.0, N, L, RS)
    assert np.abs(acc.sum(axis=0)).max() <= 1e-12 * np.abs(acc).max()


def test_uniform_lattice_feels_no_force():
    g = (np.stack(np.meshgrid(*[np.arange(N)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) * H
    acc = pm.pm_force(g, 1.0, N, L, RS)
    # compare with the field of a single particle a few cells away
    ref = np.abs(field(np.array([[0.5 * L] * 3]), 1.0, np.array([[0.5 * L + 3 * H, 0.5 * L, 0.5 * L]]))).max()
    assert np.abs(acc).max() <= 1e-9 * ref


def test_linear_in_mass_and_shift_invariant():
    rng = np.random.Generator(np.random.PCG64(2))
    pos = rng.random((200, 3)) * L
    a1 = pm.pm_force(pos, 1.0, N, L, RS)
    a3 = pm.pm_force(pos, 3.0, N, L, RS)
    assert np.allclose(a3, 3 * a1, rtol=1e-12, atol=1e-14 * np.abs(a1).max())
    shifted = np.mod(pos + np.array([H, 2 * H, -H]), L)
    a_s = pm.pm_force(shifted, 1.0, N, L, RS)
    assert np.allclose(a_s, a1, rtol=0, atol=1e-9 * np.abs(a1).max())


@pytest.mark.parametrize("v", [0, 1, 5, 16])
def test_green_function_special_values(v):
    gf = pm.green(N, L, RS)
    assert gf[0, 0, 0] == L * L / (pm.PI_TRUNC * N ** 3)
    # symmetric under axis permutation and sign of the wave vector
    assert gf[v, 0, 0] == gf[0, v, 0] == gf[0, 0, v] == gf[(-v) % N, 0, 0]
