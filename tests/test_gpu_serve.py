"""The serving loop (photons_server_*): resident boxes, continuous batching of
single-step requests, kick/drift in FP64 on the GPU.

  - one box stepped through the server equals the oracle's reference step
    (src/photoNs.c:208-268 without PM) to the FP64 force tolerance, and a box with a
    mesh the oracle's whole TreePM step (src/photoNs.c:136-268);
  - requests queued together are batched, several boxes per batch, and the lanes
    evaluate their batches at the same time;
  - each box's state after its steps is the same alone and in a mix of boxes whose
    requests arrive while others run on two overlapping lanes (FP64: round-off;
    FP32-mixed: the order of float sums);
  - admission by HBM budget refuses what does not fit.
"""
import threading

import numpy as np
import pytest

import ic
import oracle
import photons_gpu as pg
import pm

pytestmark = pytest.mark.gpu

LOGA = np.log([0.30, 0.31, 0.32])


def factors():
    out = []
    for li, lf in zip(LOGA[:-1], LOGA[1:]):
        out.append((0.5 * pg.kick_factor(li, lf) * ic.GRAV_CONST, pg.drift_factor(li, lf)))
    return out


@pytest.fixture(scope="module")
def boxes():
    counts, lengths, pos, prm = ic.workload_ensemble(4, 20 ** 3, 28 ** 3, seed=11, order=4)
    off = np.concatenate([[0], np.cumsum(counts)])
    rng = np.random.Generator(np.random.PCG64(12))
    vel = rng.normal(size=pos.shape) * 300.0
    out = [(pos[off[b]:off[b + 1]], vel[off[b]:off[b + 1]], float(lengths[b])) for b in range(len(counts))]
    return out, prm


def pdiff(a, b, L):
    d = np.abs(a - b)
    return np.minimum(d, L - d)  # positions compare modulo the box


def run_alone(prm, precision, box, nsteps):
    pos, vel, L = box
    srv = pg.Server(prm, precision, nlanes=1, hbm_budget=8e9, max_batch_particles=1 << 20)
    bid = srv.add_box(pos, vel, L)
    for dkh, dd in factors()[:nsteps]:
        srv.wait(srv.submit(bid, dkh, dd))
    out = srv.read_box(bid)
    srv.close()
    return out


def test_one_box_equals_the_reference_step(boxes):
    (box0, *_), prm = boxes
    pos, vel, L = box0
    p = ic.ForceParams(**{**prm.as_dict(), "boxsize": L})
    acc, _ = oracle.force(pos, p)
    want_p, want_v, want_a = pos, vel, acc
    for dkh, dd in factors():
        want_p, want_v, want_a = oracle.step(want_p, want_v, want_a, p, dkh, dd)
    got_p, got_v, got_a = run_alone(prm, pg.FP64, box0, 2)
    assert pdiff(got_p, want_p, L).max() <= 1e-9 * L
    assert np.abs(got_v - want_v).max() <= 1e-9 * np.abs(want_v).max()
    e = np.linalg.norm(got_a - want_a, axis=1) / np.linalg.norm(want_a, axis=1)
    assert e.max() <= 1e-9


def test_one_box_treepm_step_equals_the_oracle(boxes):
    """a box added with a mesh takes the reference's whole step: PM half kicks beside the
    short-range ones, both forces recomputed after the drift (oracle.treepm_step: the
    oracle's FMM and oracle/pm.py); the mesh is the reference's NSIDE = 1.25 L / rs"""
    (box0, *_), prm = boxes
    pos, vel, L = box0
    p = ic.ForceParams(**{**prm.as_dict(), "boxsize": L})
    nside = 2 * int(round(0.625 * L / prm.rs))
    pm_force = lambda x: pm.pm_force(x, p.mass, nside, L, p.rs)
    want = (pos, vel, oracle.force(pos, p)[0], pm_force(pos))
    for dkh, dd in factors():
        want = oracle.treepm_step(*want, p, pm_force, dkh, dd)
    srv = pg.Server(prm, pg.FP64, nlanes=1, hbm_budget=8e9, max_batch_particles=1 << 20)
    print(f"served TreePM step: n={len(pos)} nside={nside}", flush=True)
    bid = srv.add_box(pos, vel, L, nside=nside)
    for dkh, dd in factors():
        srv.wait(srv.submit(bid, dkh, dd))
    gp, gv, ga = srv.read_box(bid)
    gpm = srv.read_pm(bid)
    srv.close()
    wp, wv, wa, wpm = want
    assert pdiff(gp, wp, L).max() <= 1e-9 * L
    assert np.abs(gv - wv).max() <= 1e-9 * np.abs(wv).max()
    e = np.linalg.norm(ga - wa, axis=1) / np.linalg.norm(wa, axis=1)
    assert e.max() <= 1e-9
    assert np.abs(gpm - wpm).max() <= 1e-9 * np.abs(wpm).max()
    # the long-range part moved the particles: without it the step differs
    assert np.abs(gv - run_alone(prm, pg.FP64, box0, 2)[1]).max() > 1e-6 * np.abs(wv).max()


def test_batches_hold_several_boxes_and_lanes_overlap(boxes):
    """requests queued while the server is paused are batched when it resumes: with room
    for two boxes per batch, one lane takes two boxes and the other lane the other two,
    and the two batches run at the same time"""
    bl, prm = boxes
    # sizes a <= b <= c <= d queued as a, d, b, c with room for max(a + d, b + c): any
    # three boxes overflow it, so every batch holds two whatever the lanes' timing
    a, b, c, d = sorted(range(len(bl)), key=lambda k: len(bl[k][0]))
    order = [bl[k] for k in (a, d, b, c)]
    n = [len(x[0]) for x in order]
    cap = max(n[0] + n[1], n[2] + n[3])
    srv = pg.Server(prm, pg.FP32_MIXED, nlanes=2, hbm_budget=16e9, max_batch_particles=cap)
    srv.pause()
    ids = [srv.add_box(p, v, L) for p, v, L in order]
    rids = [srv.submit(b, *factors()[0]) for b in ids]
    srv.pause(False)
    for r in rids:
        srv.wait(r)
    srv.drain()
    st = srv.stats()
    print("batches", st["batches"], "mean boxes per batch", st["mean_boxes_per_batch"], "lanes at once",
          st["max_lanes_busy"], flush=True)
    srv.close()
    assert st["requests_done"] == 2 * len(bl)  # one first force and one step per box
    assert st["batches"] == 4 and st["mean_boxes_per_batch"] == 2.0
    assert st["max_lanes_busy"] == 2


@pytest.mark.parametrize("precision", [pg.FP64, pg.FP32_MIXED])
def test_alone_equals_in_the_mix(boxes, precision):
    bl, prm = boxes
    alone = [run_alone(prm, precision, b, 2) for b in bl]
    srv = pg.Server(prm, precision, nlanes=2, hbm_budget=16e9, max_batch_particles=1 << 20)
    ids = [srv.add_box(p, v, L) for p, v, L in bl]
    fac = factors()
    lat = []

    def client(items):
        for b, s in items:
            lat.append(srv.wait(srv.submit(ids[b], *fac[s])))

    # steps of one box must keep their order: each client owns whole boxes
    per_box = [[(b, s) for s in range(2)] for b in range(len(bl))]
    th = [threading.Thread(target=client, args=(per_box[b],)) for b in range(len(bl))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.drain()
    st = srv.stats()
    assert st["requests_done"] == 3 * len(bl)  # one initial force per box plus two steps
    assert all(x > 0 for x in lat)
    tol = 1e-12 if precision == pg.FP64 else 1e-6
    for b, (p, v, L) in enumerate(bl):
        gp, gv, ga = srv.read_box(ids[b])
        ap, av, aa = alone[b]
        assert pdiff(gp, ap, L).max() <= tol * L, b
        assert np.abs(gv - av).max() <= tol * np.abs(av).max(), b
        e = np.linalg.norm(ga - aa, axis=1) / np.linalg.norm(aa, axis=1)
        assert np.percentile(e, 99) <= (1e-12 if precision == pg.FP64 else 1e-5), b
    print("batches", st["batches"], "mean boxes per batch", st["mean_boxes_per_batch"], "lanes at once",
          st["max_lanes_busy"])
    srv.close()


def test_admission_by_hbm_budget(boxes):
    bl, prm = boxes
    p, v, L = bl[0]
    # the lanes reserve ~1 KB per particle of their capacity; the rest holds 72 B per resident particle
    cap = 1 << 16
    budget = 1024 * cap + 72 * len(p) * 1.5
    srv = pg.Server(prm, pg.FP32_MIXED, nlanes=1, hbm_budget=budget, max_batch_particles=cap)
    srv.add_box(p, v, L)
    with pytest.raises(pg.PhotonsError, match="budget"):
        srv.add_box(p, v, L)
    srv.close()
    with pytest.raises(pg.PhotonsError):
        pg.Server(prm, pg.FP32_MIXED, nlanes=2, hbm_budget=1e6, max_batch_particles=cap)
