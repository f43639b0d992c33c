"""Many independent periodic boxes in one evaluation (photons_force_boxes_*).

The forest evaluation must give every box the result it gets alone, bit for bit, in
both precisions: the trees, lists and periodic shifts never cross boxes, and the P2P
lists of a small box are cut into slices by its own particle count, not by the forest's,
so the sums run in the same order.  FP64 also equals the FP64 oracle to 1e-10.  Boxes
share the density (and so rs, rc, softening and mass) of the 256^3 workload, as in the
ensemble workload.
"""
import numpy as np
import pytest

import ic
import oracle
import photons_gpu as pg

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


@pytest.fixture(scope="module")
def ensemble():
    counts, lengths, pos, prm = ic.workload_ensemble(6, 24 ** 3, 40 ** 3, seed=5, order=4)
    off = np.concatenate([[0], np.cumsum(counts)])
    return counts, lengths, pos, prm, off


def alone(prm, L, pos, precision):
    p = ic.ForceParams(**{**prm.as_dict(), "boxsize": float(L)})
    ctx = pg.Context(p, precision)
    acc = ctx.force(pos)
    return acc, ctx.stats()


@pytest.mark.parametrize("precision", [pg.FP64, pg.FP32_MIXED])
def test_each_box_as_alone(ensemble, precision):
    counts, lengths, pos, prm, off = ensemble
    ctx = pg.Context(prm, precision)
    acc = ctx.force_boxes(counts, lengths, pos)
    st = ctx.stats()
    tot = {k: 0 for k in ("nleaf", "nnode", "p2p_pairs", "m2l_pairs", "p2p_interactions")}
    for b in range(len(counts)):
        sl = slice(off[b], off[b + 1])
        a1, s1 = alone(prm, lengths[b], pos[sl], precision)
        e = rel(acc[sl], a1)
        print(f"box {b}: n={counts[b]} L={lengths[b]:.0f} p50 {np.median(e):.1e} p99 {np.percentile(e, 99):.1e}"
              f" max {e.max():.1e}", flush=True)
        assert np.array_equal(acc[sl], a1), b
        for k in tot:
            tot[k] += s1[k]
    for k in tot:
        assert st[k] == tot[k], k


def test_forest_fp64_matches_oracle(ensemble):
    counts, lengths, pos, prm, off = ensemble
    acc = pg.Context(prm, pg.FP64).force_boxes(counts, lengths, pos)
    for b in (0, len(counts) - 1):
        sl = slice(off[b], off[b + 1])
        p = ic.ForceParams(**{**prm.as_dict(), "boxsize": float(lengths[b])})
        want, _ = oracle.force(pos[sl], p)
        assert rel(acc[sl], want).max() <= 1e-10


def test_forest_matches_reference_golden():
    """three boxes of one explicit parameter set evaluated as one forest: each box's
    force equals the reference's force of that box alone (tests/golden/forest.npz, written
    by the reference FMM subset per box) to 1e-10, and the list counters are the sums of
    the reference's per-box counters"""
    import os

    from cases import forest_inputs
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "forest.npz"),
                allow_pickle=False)
    counts, lengths, pos, prm = forest_inputs()
    np.testing.assert_array_equal(counts, g["counts"])
    np.testing.assert_array_equal(lengths, g["lengths"])
    ctx = pg.Context(prm, pg.FP64)
    acc = ctx.force_boxes(counts, lengths, pos)
    e = rel(acc[g["idx"]], g["acc"])
    print(f"forest vs reference: max {e.max():.2e}")
    assert e.max() <= 1e-10
    st = ctx.stats()
    c = g["counters"].sum(axis=0)
    assert st["p2p_local_pairs"] == c[0] and st["m2l_local_pairs"] == c[1]
    assert st["m2l_pairs"] == c[2] and st["p2p_image_interactions"] == c[3]
    assert st["nleaf"] == c[4] and st["nnode"] == c[5]


def test_fp32_forest_accuracy(ensemble):
    counts, lengths, pos, prm, off = ensemble
    a64 = pg.Context(prm, pg.FP64).force_boxes(counts, lengths, pos)
    a32 = pg.Context(prm, pg.FP32_MIXED).force_boxes(counts, lengths, pos)
    assert np.percentile(rel(a32, a64), 99) <= 1e-4


def test_bad_boxes_are_refused(ensemble):
    counts, lengths, pos, prm, off = ensemble
    ctx = pg.Context(prm, pg.FP32_MIXED)
    bad = counts.copy()
    bad[1] = 0
    with pytest.raises(pg.PhotonsError):
        ctx.force_boxes(bad, lengths, pos[: bad.sum()])
    badl = lengths.copy()
    badl[2] = -1.0
    with pytest.raises(pg.PhotonsError):
        ctx.force_boxes(counts, badl, pos)
    small = counts.copy()
    small[0] = prm.maxleaf  # periodic boxes need more than maxleaf particles
    with pytest.raises(pg.PhotonsError):
        ctx.force_boxes(small, lengths, pos[: small.sum()])
    # the context still works afterwards
    acc = ctx.force_boxes(counts[:2], lengths[:2], pos[: counts[:2].sum()])
    assert np.isfinite(acc).all()


def test_p2p_stage_variants_agree(ensemble):
    counts, lengths, pos, prm, off = ensemble
    ctx = pg.Context(prm, pg.FP32_MIXED)
    assert ctx.p2p_stage() == 256
    a256 = ctx.force_boxes(counts, lengths, pos)
    ctx.set_p2p_stage(512)
    a512 = ctx.force_boxes(counts, lengths, pos)
    assert np.percentile(rel(a256, a512), 99) <= 1e-5
    with pytest.raises(pg.PhotonsError):
        ctx.set_p2p_stage(300)


def test_autotuner_keeps_the_fastest_and_its_result(ensemble):
    import torch
    counts, lengths, pos, prm, off = ensemble
    d_pos = torch.from_numpy(np.ascontiguousarray(pos)).cuda()
    d_acc = torch.zeros_like(d_pos)
    ctx = pg.Context(prm, pg.FP32_MIXED)
    leaf, stage, table = ctx.autotune_boxes(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr(), reps=1)
    print("autotune", leaf, stage, table)
    assert (leaf, stage) in table and table[(leaf, stage)] == min(table.values())
    assert set(table) == {(lf, st) for lf in (16, 32, 64) for st in (256, 512)}
    got = d_acc.cpu().numpy()
    p = ic.ForceParams(**{**prm.as_dict(), "maxleaf": leaf})
    c2 = pg.Context(p, pg.FP32_MIXED)
    c2.set_p2p_stage(stage)
    again = c2.force_boxes(counts, lengths, pos)
    assert np.percentile(rel(got, again), 99) <= 1e-5
    a64 = pg.Context(p, pg.FP64).force_boxes(counts, lengths, pos)
    assert np.percentile(rel(got, a64), 99) <= 1e-4
