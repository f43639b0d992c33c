"""The oracle's time integration (src/initial.c:650-683, src/photoNs.c:208-268),
pinned by closed forms and by the structure of the step, not by retyping it."""
import numpy as np
import pytest
from scipy.integrate import quad

import ic
import oracle


def H(a, om, ox):
    return 0.1 * np.sqrt(om * a ** -3 + ox)


@pytest.mark.parametrize("ai,af", [(0.02, 0.025), (0.3, 0.31), (0.5, 1.0)])
def test_kick_drift_against_quadrature(ai, af):
    li, lf = np.log(ai), np.log(af)
    k = oracle.kick_loga(li, lf)
    d = oracle.drift_loga(li, lf)
    kq = quad(lambda x: 1.0 / (np.exp(x) * H(np.exp(x), 0.25, 0.75)), li, lf, epsabs=0, epsrel=1e-13)[0]
    dq = quad(lambda x: 1.0 / (np.exp(2 * x) * H(np.exp(x), 0.25, 0.75)), li, lf, epsabs=0, epsrel=1e-13)[0]
    # composite Simpson with 128 blocks: error far below 1e-9 for these smooth integrands
    assert abs(k / kq - 1) < 1e-9 and abs(d / dq - 1) < 1e-9


def test_kick_drift_einstein_de_sitter_closed_form():
    # Om = 1, OL = 0: H = 0.1 a^-3/2, kick = 20 (sqrt(af) - sqrt(ai)), drift = 20 (1/sqrt(ai) - 1/sqrt(af))
    ai, af = 0.1, 0.4
    k = oracle.kick_loga(np.log(ai), np.log(af), 1.0, 0.0)
    d = oracle.drift_loga(np.log(ai), np.log(af), 1.0, 0.0)
    assert abs(k - 20 * (np.sqrt(af) - np.sqrt(ai))) < 1e-9 * k
    assert abs(d - 20 * (1 / np.sqrt(ai) - 1 / np.sqrt(af))) < 1e-9 * d


@pytest.fixture(scope="module")
def box():
    pos, prm = ic.workload_clustered(nside=12, order=3)
    rng = np.random.Generator(np.random.PCG64(3))
    vel = rng.normal(size=pos.shape) * 300.0
    acc0, _ = oracle.force(pos, prm)
    return pos, vel, acc0, prm


def test_step_without_kick_is_a_wrapped_drift(box):
    pos, vel, acc0, prm = box
    dd = 0.37 * prm.boxsize / np.abs(vel).max()  # moves many particles across the boundary
    p1, v1, a1 = oracle.step(pos, vel, acc0, prm, 0.0, dd)
    want = np.mod(pos + vel * dd, prm.boxsize)
    assert np.allclose(p1, want, rtol=0, atol=1e-9 * prm.boxsize)
    assert (p1 >= 0).all() and (p1 < prm.boxsize).all()
    assert np.array_equal(v1, vel)
    a_new, _ = oracle.force(p1, prm)
    assert np.array_equal(a1, a_new)


def test_step_without_drift_is_two_half_kicks(box):
    pos, vel, acc0, prm = box
    dkh = 1.7e-3
    p1, v1, a1 = oracle.step(pos, vel, acc0, prm, dkh, 0.0)
    assert np.array_equal(p1, pos)
    assert np.array_equal(a1, acc0)  # same positions, same force
    assert np.allclose(v1, vel + 2 * acc0 * dkh, rtol=1e-14, atol=0)


def test_step_uses_the_old_force_then_the_new_one(box):
    pos, vel, acc0, prm = box
    dkh, dd = 1e-3, 2.0
    p1, v1, a1 = oracle.step(pos, vel, acc0, prm, dkh, dd)
    vh = vel + acc0 * dkh
    assert np.allclose(p1, np.mod(pos + vh * dd, prm.boxsize), rtol=0, atol=1e-9 * prm.boxsize)
    a_new, _ = oracle.force(p1, prm)
    assert np.allclose(v1, vh + a_new * dkh, rtol=1e-14, atol=0)


def test_treepm_step_without_pm_equals_the_c_step(box):
    # the numpy TreePM step with no long-range part is the C restatement's step, bit for bit
    pos, vel, acc0, prm = box
    dkh, dd = 1e-3, 2.0
    want = oracle.step(pos, vel, acc0, prm, dkh, dd)
    zero = np.zeros_like(pos)
    got = oracle.treepm_step(pos, vel, acc0, zero, prm, lambda p: np.zeros_like(p), dkh, dd)
    for g, w in zip(got[:3], want):
        assert np.array_equal(g, w)
    assert not got[3].any()


def test_treepm_step_puts_the_long_range_kicks_around_the_short_range_ones(box):
    # a uniform long-range field that changes between steps: the old one kicks before
    # the drift, the new one after the new short-range force
    pos, vel, acc0, prm = box
    amax = np.abs(acc0).max()
    g_old = np.array([0.5, -1.0, 0.25]) * amax
    g_new = np.array([3.0, -2.0, 1.0]) * amax
    apm0 = np.broadcast_to(g_old, pos.shape)
    dkh, dd = 1e-3, 2.0
    p1, v1, a1, apm1 = oracle.treepm_step(pos, vel, acc0, apm0, prm, lambda p: np.broadcast_to(g_new, p.shape),
                                          dkh, dd)
    vh = vel + g_old * dkh + acc0 * dkh
    assert np.allclose(p1, np.mod(pos + vh * dd, prm.boxsize), rtol=0, atol=1e-9 * prm.boxsize)
    a_new, _ = oracle.force(p1, prm)
    assert np.array_equal(a1, a_new)
    assert np.array_equal(apm1, np.broadcast_to(g_new, pos.shape))
    assert np.allclose(v1, vh + (a_new + g_new) * dkh, rtol=1e-13, atol=1e-13 * np.abs(v1).max())


def test_wrap_folds_every_coordinate_into_the_box():
    L = 7.0
    x = np.array([-20.5, -7.0, -1e-17, 0.0, 6.999999, 7.0, 13.9, 30.0])
    w = oracle.wrap(x, L)
    assert ((w >= 0) & (w < L)).all()
    d = np.abs(w - np.mod(x, L))
    assert (np.minimum(d, L - d) <= 1e-12).all()  # -1e-17 + L rounds to L, which wraps to 0
