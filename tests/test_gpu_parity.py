"""GPU short-range force against the reference (golden vectors) and the CPU oracle.

FP64 path: acceleration within 1e-10 (relative, per particle) of the reference's own
FMM and of the oracle; identical tree, interaction lists and list counters.
FP32-mixed path: 99th-percentile relative error within 1e-4 of the FP64 path.
"""
import os

import numpy as np
import pytest

import ic
import oracle
import photons_gpu as pg
from cases import FORCE_CASES, LARGE_CASES, case_inputs

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


@pytest.fixture(scope="module")
def oracle_cache():
    return {}


def oracle_force(cache, name):
    if name not in cache:
        pos, prm = case_inputs(name)
        cache[name] = oracle.force(pos, prm)
    return cache[name]


@pytest.mark.parametrize("name", sorted(FORCE_CASES))
def test_fp64_matches_reference_golden(name, oracle_cache):
    print(f"case {name}", flush=True)
    g = np.load(os.path.join(GOLDEN, f"force_{name}.npz"), allow_pickle=False)
    pos, prm = case_inputs(name)
    ctx = pg.Context(prm, pg.FP64)
    acc = ctx.force(pos)
    assert np.isfinite(acc).all()
    e = rel_err(acc[g["idx"]], g["acc"])
    assert e.max() <= 1e-10, (name, e.max())
    st = ctx.stats()
    c = g["counters"]
    assert st["p2p_local_pairs"] == c[0]
    assert st["m2l_local_pairs"] == c[1]
    assert st["m2l_pairs"] == c[2]
    assert st["p2p_image_interactions"] == c[3]
    assert st["nleaf"] == c[4] and st["nnode"] == c[5]
    print(f"{name}: {len(pos)} particles, max relative difference to the reference {e.max():.2e}")
    if name in LARGE_CASES:
        return  # the stored vectors are the reference here; the oracle would take a minute
    ref, _ = oracle_force(oracle_cache, name)
    e = rel_err(acc, ref)
    assert e.max() <= 1e-10, (name, e.max())


@pytest.mark.parametrize("name", ["clus24_pm48", "clus32_pm32", "lattice32_rc6"])
def test_tree_and_lists_match_oracle(name):
    pos, prm = case_inputs(name)
    ctx = pg.Context(prm, pg.FP64)
    ctx.force(pos)
    t = oracle.Tree(pos, prm)
    ex = t.export()
    cells = ctx.cells()
    nl = t.nleaf
    assert cells["nleaf"] == nl and cells["nnode"] == t.nnode
    np.testing.assert_array_equal(cells["npart"][:nl], ex["leaf_npart"])
    np.testing.assert_array_equal(cells["ipart"][:nl], ex["leaf_ipart"])
    np.testing.assert_array_equal(cells["npart"][nl:], ex["node_npart"])
    osons = np.where(ex["node_son"] < 0, -ex["node_son"] - 1, ex["node_son"] + nl)
    np.testing.assert_array_equal(cells["son"][nl:], osons)
    ocent = np.concatenate([ex["leaf_center"], ex["node_center"]])
    owid = np.concatenate([ex["leaf_width"], ex["node_width"]])
    assert np.abs(cells["center"] - ocent).max() <= 1e-12 * prm.boxsize
    assert np.abs(cells["width"] - owid).max() <= 1e-12 * prm.boxsize
    # same particles in every leaf
    perm = ctx.perm(len(pos))
    for l in range(0, nl, max(1, nl // 200)):
        a, b = ex["leaf_ipart"][l], ex["leaf_ipart"][l] + ex["leaf_npart"][l]
        assert set(perm[a:b]) == set(ex["perm"][a:b])
    # lists: identical entries, and per target the reference's order
    for got, want in ((ctx.p2p_list(), t.p2p_list()), (ctx.m2l_list(), t.m2l_list())):
        gk = np.stack([got[0], got[1], got[2].astype(np.int32)], axis=1)
        wk = np.stack([want[0], want[1], want[2].astype(np.int32)], axis=1)
        order = np.argsort(wk[:, 0], kind="stable")
        np.testing.assert_array_equal(gk, wk[order])


@pytest.mark.parametrize("name", ["clus32_pm32", "lcdm_demo", "clus24_pm48_th0.3"])
def test_fp32_mixed_error(name):
    pos, prm = case_inputs(name)
    a64 = pg.Context(prm, pg.FP64).force(pos)
    a32 = pg.Context(prm, pg.FP32_MIXED).force(pos)
    e = rel_err(a32, a64)
    print(f"{name}: fp32-mixed vs fp64 p50 {np.median(e):.2e} p99 {np.percentile(e, 99):.2e} max {e.max():.2e}")
    assert np.percentile(e, 99) <= 1e-4
    assert np.median(e) <= 1e-5


@pytest.mark.parametrize("order", [4, 5])
def test_higher_orders_match_oracle(order):
    pos, prm = case_inputs("clus24_pm48")
    prm.order = order
    ref, st = oracle.force(pos, prm)
    assert st["m2l_calls"] > 0
    acc = pg.Context(prm, pg.FP64).force(pos)
    assert rel_err(acc, ref).max() <= 1e-10
    a32 = pg.Context(prm, pg.FP32_MIXED).force(pos)
    assert np.percentile(rel_err(a32, ref), 99) <= 1e-4


@pytest.mark.parametrize("maxleaf", [2, 16, 64, 100, 256])
def test_leaf_sizes_match_oracle(maxleaf):
    """the reference takes the leaf size from its parameter file (MaxPackage,
    src/initial.c:172-173); leaves of more than 64 particles run in chunks of 64 targets"""
    pos, prm = case_inputs("clus16_pm32")
    prm.maxleaf = maxleaf
    ref, _ = oracle.force(pos, prm)
    acc = pg.Context(prm, pg.FP64).force(pos)
    assert rel_err(acc, ref).max() <= 1e-10
    a32 = pg.Context(prm, pg.FP32_MIXED).force(pos)
    assert np.percentile(rel_err(a32, ref), 99) <= 1e-4


def test_leaf_size_above_the_limit_is_refused():
    _, prm = case_inputs("clus16_pm32")
    prm.maxleaf = 257
    with pytest.raises(pg.PhotonsError, match="maxleaf"):
        pg.Context(prm, pg.FP64)


def test_repeatable_bit_for_bit():
    pos, prm = case_inputs("clus32_pm32")
    for prec in (pg.FP64, pg.FP32_MIXED):
        ctx = pg.Context(prm, prec)
        a = ctx.force(pos)
        b = ctx.force(pos)
        c = pg.Context(prm, prec).force(pos)
        assert np.array_equal(a, b) and np.array_equal(a, c)


def test_context_reuse_across_sizes():
    """A context keeps its buffers; a smaller and then larger input give fresh results."""
    ctx = pg.Context(case_inputs("clus32_pm32")[1], pg.FP64)
    for name in ("clus32_pm32", "clus16_pm32", "lattice32_rc6"):
        pos, prm = case_inputs(name)
        ctx.set_params(prm)
        ref, _ = oracle.force(pos, prm)
        assert rel_err(ctx.force(pos), ref).max() <= 1e-10


def _plane_value(n, L):
    """a coordinate v whose mean over n copies rounds below v, both in the reference's
    sequential sum and in the device's chunked sum (64 lanes, then a butterfly)"""
    for k in range(1, 2000):
        v = L * k / 2003.0
        seq = 0.0
        for _ in range(n):
            seq += v
        lanes = np.zeros(64)
        for i in range(n):
            lanes[i % 64] += v
        o = 32
        while o >= 1:
            lanes = lanes + lanes[np.arange(64) ^ o]
            o //= 2
        if seq / n < v and lanes[0] / n < v:
            return v
    raise AssertionError("no such coordinate")


def test_plane_of_equal_coordinates_matches_oracle():
    """100 particles sharing one x coordinate whose mean rounds below it: the reference's
    two-pointer partition (src/fmm.c:56-77) leaves the root's left child empty and passes
    every particle right; tree, lists and force are the oracle's"""
    _, prm = case_inputs("clus16_pm32")
    n = 100
    rng = np.random.default_rng(17)
    pos = rng.uniform(0.0, prm.boxsize, (n, 3))
    pos[:, 0] = _plane_value(n, prm.boxsize)
    ctx = pg.Context(prm, pg.FP64)
    acc = ctx.force(pos)
    t = oracle.Tree(pos, prm)
    ex = t.export()
    cells = ctx.cells()
    nl = t.nleaf
    assert cells["nleaf"] == nl and cells["nnode"] == t.nnode
    root = nl  # nodes follow the leaves, the root first
    left = cells["son"][root][0]
    assert left < nl and cells["npart"][left] == 0
    np.testing.assert_array_equal(cells["npart"][:nl], ex["leaf_npart"])
    np.testing.assert_array_equal(cells["npart"][nl:], ex["node_npart"])
    ref, _ = oracle.force(pos, prm)
    assert rel_err(acc, ref).max() <= 1e-10


def test_zero_softening_matches_oracle():
    """SOFTENING 0 is a valid reference setting (src/initial.c:123 refuses only < 0): a
    particle with itself then adds nothing, as the reference's skip of that pair"""
    pos, prm = case_inputs("clus16_pm32")
    prm.soft = 0.0
    ref, _ = oracle.force(pos, prm)
    assert np.isfinite(ref).all()
    a64 = pg.Context(prm, pg.FP64).force(pos)
    assert np.isfinite(a64).all()
    assert rel_err(a64, ref).max() <= 1e-10
    a32 = pg.Context(prm, pg.FP32_MIXED).force(pos)
    assert np.isfinite(a32).all()
    assert np.percentile(rel_err(a32, ref), 99) <= 1e-4


def test_edge_inputs():
    _, prm = case_inputs("clus16_pm32")
    ctx = pg.Context(prm, pg.FP64)
    assert ctx.force(np.zeros((0, 3))).shape == (0, 3)
    with pytest.raises(pg.PhotonsError):  # periodic needs more than maxleaf particles
        ctx.force(np.random.default_rng(0).uniform(0, prm.boxsize, (prm.maxleaf, 3)))
    # non-periodic small inputs, including a single leaf under the root
    for n in (1, 2, 3, 33, 100):
        pos = np.random.default_rng(n).uniform(0, prm.boxsize, (n, 3))
        p = prm.as_dict()
        p["periodic"] = 0
        ref, _ = oracle.force(pos, p)
        acc = pg.Context(p, pg.FP64).force(pos)
        scale = np.abs(ref).max() + 1e-300
        assert np.abs(acc - ref).max() <= 1e-10 * scale, n
    # more than maxleaf particles at one position: where the reference's partition would
    # recurse forever (every mean rounding low), the node splits off one particle
    rng = np.random.default_rng(5)
    pos = np.concatenate([rng.uniform(0, prm.boxsize, (300, 3)), np.repeat(rng.uniform(0, prm.boxsize, (1, 3)), 50, 0)])
    ctx = pg.Context(prm, pg.FP64)
    acc = ctx.force(pos)
    assert np.isfinite(acc).all()
    assert ctx.stats()["depth"] >= (50 - prm.maxleaf)
    # the coincident particles feel nearly the same force: they sit in different leaves,
    # so pairs near the cutoff can differ, and the split kernel is small there
    dup = acc[300:]
    assert np.abs(dup - dup[0]).max() <= 1e-4 * np.abs(dup[0]).max()
    assert np.array_equal(acc, ctx.force(pos))
