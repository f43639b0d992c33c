"""The CPU oracle against the golden vectors written by the compiled reference.

The fixtures under tests/golden/ were produced by oracle/_ref/refharness, which runs
the reference's own src/fmm.c, src/operator.c, src/remotes.c and src/toptree.c on one
MPI rank (tests/golden/make_golden.py).
"""
import hashlib
import os

import numpy as np
import pytest

import oracle
from cases import FORCE_CASES, LARGE_CASES, case_inputs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.mark.parametrize("name", sorted(FORCE_CASES))
def test_force_matches_reference(name, full_sizes):
    if name in LARGE_CASES and not full_sizes:
        pytest.skip("a minute of oracle time: run with --full")
    g = load(f"force_{name}.npz")
    pos, prm = case_inputs(name)
    digest = hashlib.sha256(np.ascontiguousarray(pos, np.float64).tobytes()).hexdigest()
    assert digest == str(g["pos_sha256"]), "input generator drifted from the golden inputs"
    acc, st = oracle.force(pos, prm)
    idx = g["idx"]
    ref = g["acc"]
    rel = np.linalg.norm(acc[idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rel.max() < 1e-12, rel.max()
    np.testing.assert_allclose(np.abs(acc).sum(axis=0), g["acc_abs_sum"], rtol=1e-12)
    np.testing.assert_allclose((acc * acc).sum(axis=0), g["acc_sq_sum"], rtol=1e-12)
    scale = g["acc_abs_sum"].sum()
    assert np.abs(acc.sum(axis=0) - g["acc_sum"]).max() < 1e-12 * scale
    c = g["counters"]
    assert st["p2p_local_pairs"] == c[0]  # idxP2P
    assert st["m2l_local_pairs"] == c[1]  # idxM2L
    assert st["m2l_calls"] == c[2]  # walk_m2l_count
    assert st["p2p_ext_inter"] == c[3]  # p2p_count_remote
    assert st["nleaf"] == c[4] and st["nnode"] == c[5]


def _split(arr, counts):
    return np.split(arr, np.cumsum(counts)[:-1])


def test_p2m_matches_reference():
    g = load("ops.npz")
    pos = _split(g["p2m_pos"], g["p2m_n"])
    for k, p in enumerate(pos):
        M = oracle.p2m(3, p, g["p2m_mass"][k], g["p2m_center"][k])
        scale = np.abs(g["p2m_M"][k]).max(initial=0.0) + 1e-300
        np.testing.assert_allclose(M, g["p2m_M"][k], rtol=1e-13, atol=1e-15 * scale)


@pytest.mark.parametrize("op", ["m2m", "l2l", "m2l"])
def test_shift_ops_match_reference(op):
    g = load("ops.npz")
    for k in range(len(g[f"{op}_d"])):
        d, X = g[f"{op}_d"][k], g[f"{op}_in"][k]
        if op == "m2m":
            out = oracle.m2m(3, d, X)
        elif op == "l2l":
            out = oracle.l2l(3, d, X)
        else:
            out = oracle.m2l(3, g["m2l_rs"][k], d, X)
        ref = g[f"{op}_out"][k]
        # compare per order: entries of one order share a scale
        for lo, hi in ((0, 1), (1, 4), (4, 10), (10, 20)):
            scale = np.abs(ref[lo:hi]).max()
            assert np.abs(out[lo:hi] - ref[lo:hi]).max() <= 1e-13 * scale, (op, k, lo)


def test_l2p_matches_reference():
    g = load("ops.npz")
    pos = _split(g["l2p_pos"], g["l2p_n"])
    ref = _split(g["l2p_acc"], g["l2p_n"])
    for k, p in enumerate(pos):
        a = oracle.l2p(3, g["l2p_L"][k], g["l2p_center"][k], p)
        np.testing.assert_allclose(a, ref[k], rtol=1e-12, atol=1e-14 * np.abs(ref[k]).max())


def test_acceptance_matches_reference():
    g = load("ops.npz")
    for row, flag in zip(g["acc_in"], g["acc_flag"]):
        theta, rc, wi, wj, dist = row[0], row[1], row[2:5], row[5:8], row[8:11]
        assert oracle.acceptance(wi, wj, dist, theta, rc) == flag
    # every outcome is represented
    assert set(np.unique(g["acc_flag"])) == {-1, 0, 1}


def test_forest_boxes_match_reference():
    """the forest fixture: each box's force from the reference alone, reproduced by the
    oracle on that box with the shared parameter set"""
    from cases import forest_box_params, forest_inputs
    g = load("forest.npz")
    counts, lengths, pos, prm = forest_inputs()
    digest = hashlib.sha256(np.ascontiguousarray(pos, np.float64).tobytes()).hexdigest()
    assert digest == str(g["pos_sha256"]), "input generator drifted from the golden inputs"
    off = np.concatenate([[0], np.cumsum(counts)])
    idx = g["idx"]
    for b in range(len(counts)):
        acc, st = oracle.force(pos[off[b]:off[b + 1]], forest_box_params(prm, lengths[b]))
        sel = (idx >= off[b]) & (idx < off[b + 1])
        ref = g["acc"][sel]
        rel = np.linalg.norm(acc[idx[sel] - off[b]] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        assert rel.max() < 1e-12, (b, rel.max())
        c = g["counters"][b]
        assert st["p2p_local_pairs"] == c[0] and st["m2l_local_pairs"] == c[1]
        assert st["m2l_calls"] == c[2] and st["p2p_ext_inter"] == c[3]
        assert st["nleaf"] == c[4] and st["nnode"] == c[5]
