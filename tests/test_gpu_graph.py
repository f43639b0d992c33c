"""The planned, captured force evaluation (photons_graph_*).

One evaluation of a fixed set of boxes runs without a host round trip, an allocation or
a memset and is captured into a hipGraph.  Every replay must give, bit for bit, what an
ordinary evaluation of the same positions gives (the kernels are the same, only the launch
bounds differ), also after the positions changed in place between replays, and the FP64
result must match the oracle.  Particles that move so far that the tree or the lists
outgrow the plan must be reported, not computed wrongly in silence.
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
def boxes():
    counts, lengths, pos, prm = ic.workload_ensemble(4, 20 ** 3, 32 ** 3, seed=9, order=4)
    return counts, lengths, pos, prm


def device_arrays(pos):
    import torch
    d_pos = torch.from_numpy(np.ascontiguousarray(pos)).cuda()
    d_acc = torch.zeros_like(d_pos)
    return d_pos, d_acc


@pytest.mark.parametrize("precision", [pg.FP64, pg.FP32_MIXED])
@pytest.mark.parametrize("capture", [True, False])
def test_replays_equal_the_ordinary_evaluation(boxes, precision, capture):
    import torch
    counts, lengths, pos, prm = boxes
    want = pg.Context(prm, precision).force_boxes(counts, lengths, pos)
    d_pos, d_acc = device_arrays(pos)
    ctx = pg.Context(prm, precision)
    ctx.graph_create(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr(), margin=0.25, capture=capture)
    torch.cuda.synchronize()
    assert np.array_equal(d_acc.cpu().numpy(), want)
    for rep in range(3):
        d_acc.zero_()
        torch.cuda.synchronize()
        ctx.graph_launch()
        ovf, _ = ctx.graph_check()
        assert ovf == 0
        got = d_acc.cpu().numpy()
        print(f"replay {rep}: max |diff| {np.abs(got - want).max():.1e}", flush=True)
        assert np.array_equal(got, want), rep
    st = ctx.stats()
    st_ord = pg.Context(prm, precision)
    st_ord.force_boxes(counts, lengths, pos)
    for k in ("nleaf", "nnode", "p2p_pairs", "m2l_pairs", "p2p_interactions"):
        assert st[k] == st_ord.stats()[k], k
    ctx.graph_destroy()


def test_replay_follows_positions_changed_in_place(boxes):
    """a drift between replays: each replay equals an ordinary evaluation of the new
    positions, and the FP64 result matches the oracle"""
    import torch
    counts, lengths, pos, prm = boxes
    d_pos, d_acc = device_arrays(pos)
    ctx = pg.Context(prm, pg.FP64)
    ctx.graph_create(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr(), margin=0.25)
    off = np.concatenate([[0], np.cumsum(counts)])
    rng = np.random.default_rng(3)
    cur = pos.copy()
    for step in range(2):
        # move every particle by up to 1% of the mean spacing, wrapped into its box
        for b in range(len(counts)):
            sl = slice(off[b], off[b + 1])
            L = lengths[b]
            cur[sl] = np.mod(cur[sl] + rng.uniform(-0.01, 0.01, (counts[b], 3)) * L / counts[b] ** (1 / 3), L)
        d_pos.copy_(torch.from_numpy(cur))
        torch.cuda.synchronize()
        ctx.graph_launch()
        ovf, _ = ctx.graph_check()
        assert ovf == 0
        got = d_acc.cpu().numpy()
        want = pg.Context(prm, pg.FP64).force_boxes(counts, lengths, cur)
        assert np.array_equal(got, want), step
    b = 0
    p = ic.ForceParams(**{**prm.as_dict(), "boxsize": float(lengths[b])})
    ref, _ = oracle.force(cur[off[b]:off[b + 1]], p)
    assert rel(got[off[b]:off[b + 1]], ref).max() <= 1e-10


def test_outgrown_plan_is_reported(boxes):
    """particles gathered into a small corner of each box: the tree and the lists outgrow
    a plan made without margin; the replay reports it, and a new plan computes them"""
    import torch
    counts, lengths, pos, prm = boxes
    d_pos, d_acc = device_arrays(pos)
    ctx = pg.Context(prm, pg.FP32_MIXED)
    ctx.graph_create(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr(), margin=0.0)
    off = np.concatenate([[0], np.cumsum(counts)])
    dense = pos.copy()
    for b in range(len(counts)):
        sl = slice(off[b], off[b + 1])
        dense[sl] = pos[sl] * 0.15  # the same particles in 0.3% of the volume
    d_pos.copy_(torch.from_numpy(dense))
    torch.cuda.synchronize()
    ctx.graph_launch()
    ovf, _ = ctx.graph_check()
    print("overflow bits", ovf)
    assert ovf != 0
    with pytest.raises(pg.PhotonsError):
        ctx.force_boxes(counts, lengths, dense)  # held by the plan
    ctx.graph_create(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr(), margin=0.25)
    want = pg.Context(prm, pg.FP32_MIXED).force_boxes(counts, lengths, dense)
    assert np.array_equal(d_acc.cpu().numpy(), want)
    ctx.graph_launch()
    assert ctx.graph_check()[0] == 0
    assert np.array_equal(d_acc.cpu().numpy(), want)
    ctx.graph_destroy()
    assert np.array_equal(ctx.force_boxes(counts, lengths, dense), want)


@pytest.mark.parametrize("precision", [pg.FP64, pg.FP32_MIXED])
def test_reused_plan_outgrown_by_the_tree(boxes, precision):
    """Repeated evaluations of one box set run on the first one's plan (plan reuse); a
    tree that outgrows it is measured again.  Clustered positions planned, then uniform
    ones (more nodes per level than the plan holds), then the clustered ones under the
    uniform plan (more levels): both trees overflow in the first planned evaluation of
    their plan, whose buffers are new, and every result equals a fresh context's, bit for
    bit (the tree and walk must read nothing of a tree that overflowed)."""
    counts, lengths, pos, prm = boxes
    rng = np.random.default_rng(3)
    uni = np.concatenate([rng.uniform(0.0, lengths[b], (counts[b], 3)) for b in range(len(counts))])
    ctx = pg.Context(prm, precision)
    for k, p in enumerate((pos, uni, pos, pos)):
        print(f"evaluation {k}: precision {precision}", flush=True)
        got = ctx.force_boxes(counts, lengths, p)
        fresh = pg.Context(prm, precision)
        want = fresh.force_boxes(counts, lengths, p)
        assert np.array_equal(got, want), k
        for key in ("nleaf", "nnode", "p2p_pairs", "m2l_pairs"):
            assert ctx.stats()[key] == fresh.stats()[key], (k, key)


def test_small_box_latency():
    """one 16^3 clustered box: the replayed graph against the ordinary evaluation
    (host round trips per tree level and walk depth)"""
    import time

    import torch
    n = 16 ** 3
    box = 500000.0 * (n / 256.0 ** 3) ** (1.0 / 3.0)
    pos, _ = ic.clustered(n, box, seed=1)
    prm = ic.force_params(n, box, 16, rc_factor=6.0, order=4)
    d_pos, d_acc = device_arrays(pos)
    ctx = pg.Context(prm, pg.FP32_MIXED)
    counts, lengths = np.array([n]), np.array([box])
    ctx.force_boxes_device(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr())
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(20):
        ctx.force_boxes_device(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr())
    ctx.sync()
    t_ord = (time.perf_counter() - t0) / 20 * 1e3
    want = d_acc.cpu().numpy()
    g = pg.Context(prm, pg.FP32_MIXED)
    g.graph_create(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr())
    g.graph_launch()
    g.graph_check()
    t0 = time.perf_counter()
    for _ in range(20):
        g.graph_launch()
    g.graph_check()
    t_graph = (time.perf_counter() - t0) / 20 * 1e3
    print(f"16^3 box: ordinary {t_ord:.3f} ms, replayed graph {t_graph:.3f} ms per evaluation")
    assert np.array_equal(d_acc.cpu().numpy(), want)
    assert t_graph < t_ord
