"""Parity at the headline size: 256^3 clustered particles (p = 4, theta 0.4, leaf 32,
rc = 6 rs), the workload the benchmark measures.

The oracle builds the whole 256^3 tree and every multipole but walks and evaluates only
a seeded sample of target leaves (oracle.force_leaves, pinned against the full oracle
bit for bit): the GPU FP64 path must match it to 1e-10 on every particle of those leaves,
with the same P2P list per target leaf and the same number of M2L entries on the leaf
and its ancestors; and the whole P2P and M2L lists must hold as many entries as the
oracle's full walk counts (a committed fixture).  The FP32-mixed path is compared with it too (the bench's gate:
p99 <= 1e-4), and its largest relative error is reported with the particle's
acceleration against the median, since relative errors blow up where the force cancels.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import ic
import oracle
import photons_gpu as pg

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def headline():
    pos, prm = ic.workload_clustered(nside=256, order=4)
    t = oracle.Tree(pos, prm)
    ex = t.export()
    # 96 leaves: most at random, plus the most populated and some empty ones excluded
    rng = np.random.default_rng(2024)
    filled = np.flatnonzero(ex["leaf_npart"] > 0)
    sel = np.sort(rng.choice(filled, 96, replace=False)).astype(np.int32)
    acc, ids, plen, mlen = oracle.force_leaves(pos, prm, sel)
    del t
    return pos, prm, sel, acc, ids, plen, mlen


def rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def test_fp64_matches_the_oracle_at_256cubed(headline):
    pos, prm, sel, want, ids, plen, mlen = headline
    print(f"256^3: {len(sel)} target leaves, {len(ids)} particles", flush=True)
    ctx = pg.Context(prm, pg.FP64)
    acc = ctx.force(pos)
    e = rel(acc[ids], want)
    print(f"FP64 vs oracle: max {e.max():.2e}", flush=True)
    assert e.max() <= 1e-10
    # the whole lists: as many entries as the oracle's full walk counts
    # (tests/golden/make_counts256.py)
    with open(os.path.join(GOLDEN, "lists_clustered256.json")) as f:
        tot = json.load(f)
    assert hashlib.sha256(np.ascontiguousarray(pos).tobytes()).hexdigest() == tot["positions_sha256"]
    st = ctx.stats()
    assert (st["nleaf"], st["nnode"]) == (tot["nleaf"], tot["nnode"])
    assert (st["p2p_pairs"], st["m2l_pairs"]) == (tot["p2p_entries"], tot["m2l_entries"])
    # the same lists for the sampled targets
    cells = ctx.cells()
    nleaf = cells["nleaf"]
    np.testing.assert_array_equal(ctx.list_lengths("p2p")[sel], plen)
    per_cell = ctx.list_lengths("m2l")
    parent = np.full(len(per_cell), -1)
    nodes = np.arange(nleaf, len(per_cell))
    parent[cells["son"][nleaf:, 0]] = nodes
    parent[cells["son"][nleaf:, 1]] = nodes
    for k, leaf in enumerate(sel):
        tot, c = 0, int(leaf)
        while c >= 0:
            tot += per_cell[c]
            c = parent[c]
        assert tot == mlen[k], (leaf, tot, mlen[k])


def test_fp32_mixed_against_the_oracle_at_256cubed(headline):
    pos, prm, sel, want, ids, plen, mlen = headline
    acc = pg.Context(prm, pg.FP32_MIXED).force(pos)[ids]
    e = rel(acc, want)
    mag = np.linalg.norm(want, axis=1)
    k = int(np.argmax(e))
    absd = np.linalg.norm(acc[k] - want[k])
    print(f"FP32-mixed vs FP64 oracle: p50 {np.median(e):.2e} p99 {np.percentile(e, 99):.2e} max {e.max():.2e}; "
          f"worst particle |a| = {mag[k] / np.median(mag):.2e} of the median, its error "
          f"{absd / np.median(mag):.2e} of the median |a|", flush=True)
    assert np.percentile(e, 99) <= 1e-4
