"""Write the golden vectors of the short-range force path from the compiled reference.

Runs in the build container only (needs /root/reference and oracle/_ref/refharness,
built by `make -C oracle ref`).  The reference FMM subset (src/fmm.c, operator.c,
remotes.c, toptree.c, utility.c) is run by oracle/ref_harness.c on one MPI rank.

Outputs, all committed as small fixtures next to this script:
  force_<case>.npz  inputs are regenerated from seeds (a SHA-256 of the positions is
                    stored to catch generator drift); stored: acceleration of every
                    8th particle (64th for the 64^3 case), full-vector checksums, list
                    counters; written only when the oracle agrees with the reference.
  lcdm_pos.npy      positions of demo/ic_lcdm.gdt2 (float32, exactly as stored there)
  ops.npz           single-operator calls: inputs and outputs
  tree_<case>.npz   the multipoles of every node and leaf after fmm_prepare (walk_m2m)
  forest.npz        three boxes of one parameter set, each box's force alone

    python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import hashlib
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "photons-2.0_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import ic  # noqa: E402
import oracle  # noqa: E402
from cases import FORCE_CASES, STRIDE, TREE_CASES, case_inputs, forest_box_params, forest_inputs  # noqa: E402

REFERENCE = "/root/reference"


def pos_digest(pos: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(pos, np.float64).tobytes()).hexdigest()


def read_gadget_positions(path: str) -> np.ndarray:
    """Gadget-2 snapshot: 256-byte header block, then the float32 position block
    (src/snapshot.c:5-22, 211-293)."""
    raw = open(path, "rb").read()
    (h,) = struct.unpack_from("<i", raw, 0)
    npart = struct.unpack_from("<6i", raw, 4)
    n = sum(npart)
    off = 4 + h + 4
    (blk,) = struct.unpack_from("<i", raw, off)
    assert blk == 12 * n, (blk, n)
    return np.frombuffer(raw, dtype="<f4", count=3 * n, offset=off + 4).reshape(n, 3).copy()


def write_force_cases(tmp: str, names):
    lcdm = read_gadget_positions(os.path.join(REFERENCE, "demo", "ic_lcdm.gdt2"))
    np.save(os.path.join(HERE, "lcdm_pos.npy"), lcdm)
    for name in names:
        pos, prm = case_inputs(name)
        tree = oracle.Tree(pos, prm)
        if not tree.fits_reference_buffers():
            raise SystemExit(f"{name}: exceeds the reference's fixed buffers")
        acc, cnt = oracle.ref_force(pos, prm, tmp)
        oacc, _ = oracle.force(pos, prm)
        agree = float(np.max(np.linalg.norm(oacc - acc, axis=1) / np.linalg.norm(acc, axis=1)))
        if agree > 1e-12:
            raise SystemExit(f"{name}: the oracle and the reference differ by {agree:.2e}")
        idx = np.arange(0, len(pos), STRIDE.get(name, 8))
        np.savez_compressed(
            os.path.join(HERE, f"force_{name}.npz"),
            pos_sha256=np.array(pos_digest(pos)),
            params=np.array([prm.boxsize, prm.mass, prm.soft, prm.rs, prm.rc, prm.theta,
                             prm.maxleaf, prm.order]),
            idx=idx.astype(np.int32), acc=acc[idx],
            acc_sum=acc.sum(axis=0), acc_abs_sum=np.abs(acc).sum(axis=0),
            acc_sq_sum=(acc * acc).sum(axis=0),
            counters=np.array([cnt[k] for k in ("idxP2P", "idxM2L", "walk_m2l_count",
                                                "p2p_count_remote", "nleaf", "nnode")], np.int64))
        print(name, len(pos), cnt, f"oracle vs reference {agree:.1e}", flush=True)


def run_ops(records: bytes, count: int, tmp: str) -> bytes:
    fin, fout = os.path.join(tmp, "ops_in.bin"), os.path.join(tmp, "ops_out.bin")
    with open(fin, "wb") as f:
        f.write(struct.pack("<i", count))
        f.write(records)
    subprocess.run([oracle.REF_HARNESS, "ops", fin, fout], check=True)
    return open(fout, "rb").read()


def write_ops(tmp: str):
    rng = np.random.Generator(np.random.PCG64(12345))
    nt = 20
    out = {}
    # p2m: a cloud of particles around a centre
    recs, p2m_in = b"", []
    for k in range(8):
        np_ = int(rng.integers(1, 33))
        c = rng.uniform(-1e4, 1e4, 3)
        w = rng.uniform(50.0, 3000.0)
        pos = c + rng.uniform(-0.5, 0.5, (np_, 3)) * w
        mass = float(rng.uniform(10.0, 300.0))
        recs += struct.pack("<ii", 0, np_) + struct.pack("<d", mass) + c.tobytes() + pos.tobytes()
        p2m_in.append((pos, mass, c))
    raw = run_ops(recs, len(p2m_in), tmp)
    out["p2m_M"] = np.frombuffer(raw, np.float64).reshape(-1, nt)
    out["p2m_pos"] = np.concatenate([p for p, _, _ in p2m_in])
    out["p2m_n"] = np.array([len(p) for p, _, _ in p2m_in], np.int32)
    out["p2m_mass"] = np.array([m for _, m, _ in p2m_in])
    out["p2m_center"] = np.array([c for _, _, c in p2m_in])
    # m2m / l2l / m2l: random expansions and displacements
    for kind, key in ((1, "m2m"), (3, "l2l"), (2, "m2l")):
        recs, ins = b"", []
        for k in range(16):
            d = rng.uniform(-1.0, 1.0, 3) * rng.uniform(100.0, 5000.0)
            if kind == 2:
                d = rng.normal(size=3)
                d *= rng.uniform(500.0, 15000.0) / np.linalg.norm(d)
            M = rng.normal(size=nt) * np.array([1e3] + [1e5] * 3 + [1e7] * 6 + [1e9] * 10)
            if kind == 3:
                M = rng.normal(size=nt) * np.array([1e-3] + [1e-6] * 3 + [1e-9] * 6 + [1e-12] * 10)
            rs = float(rng.uniform(1000.0, 3000.0))
            if kind == 2:
                recs += struct.pack("<i", kind) + struct.pack("<d", rs) + d.tobytes() + M.tobytes()
            else:
                recs += struct.pack("<i", kind) + d.tobytes() + M.tobytes()
            ins.append((d, M, rs))
        raw = run_ops(recs, len(ins), tmp)
        out[f"{key}_out"] = np.frombuffer(raw, np.float64).reshape(-1, nt)
        out[f"{key}_d"] = np.array([d for d, _, _ in ins])
        out[f"{key}_in"] = np.array([M for _, M, _ in ins])
        out[f"{key}_rs"] = np.array([rs for _, _, rs in ins])
    # l2p
    recs, ins = b"", []
    for k in range(8):
        np_ = int(rng.integers(1, 33))
        c = rng.uniform(-1e4, 1e4, 3)
        w = rng.uniform(50.0, 3000.0)
        pos = c + rng.uniform(-0.5, 0.5, (np_, 3)) * w
        L = rng.normal(size=nt) * np.array([1e-3] + [1e-6] * 3 + [1e-9] * 6 + [1e-12] * 10)
        recs += struct.pack("<ii", 4, np_) + c.tobytes() + L.tobytes() + pos.tobytes()
        ins.append((pos, L, c))
    raw = run_ops(recs, len(ins), tmp)
    out["l2p_acc"] = np.frombuffer(raw, np.float64).reshape(-1, 3)
    out["l2p_pos"] = np.concatenate([p for p, _, _ in ins])
    out["l2p_n"] = np.array([len(p) for p, _, _ in ins], np.int32)
    out["l2p_L"] = np.array([L for _, L, _ in ins])
    out["l2p_center"] = np.array([c for _, _, c in ins])
    # acceptance: random boxes plus constructed edge cases
    recs, ins = b"", []
    for k in range(400):
        theta = float(rng.choice([0.3, 0.4, 0.5]))
        rc = float(rng.uniform(2000.0, 15000.0))
        wi = np.exp(rng.uniform(np.log(10.0), np.log(6000.0))) * rng.uniform(0.5, 1.0, 3)
        wj = np.exp(rng.uniform(np.log(10.0), np.log(6000.0))) * rng.uniform(0.5, 1.0, 3)
        dist = rng.normal(size=3)
        dist *= rng.uniform(0.0, 2.5 * rc) / np.linalg.norm(dist)
        if k % 5 == 0:  # touching along one axis
            ax = int(rng.integers(0, 3))
            dist[ax] = np.sign(dist[ax] or 1.0) * 0.5 * (wi[ax] + wj[ax])
        ins.append((theta, rc, wi, wj, dist))
    for theta, rc, wi, wj, dist in ins:
        recs += struct.pack("<i", 5) + np.array([theta, rc]).tobytes() + wi.tobytes() + wj.tobytes() + dist.tobytes()
    raw = run_ops(recs, len(ins), tmp)
    out["acc_flag"] = np.frombuffer(raw, np.int32)
    out["acc_in"] = np.array([np.concatenate([[t, r], a, b, c]) for t, r, a, b, c in ins])
    np.savez_compressed(os.path.join(HERE, "ops.npz"), **out)
    print("ops", {k: v.shape for k, v in out.items()})


def write_tree_cases(tmp: str):
    for name in TREE_CASES:
        pos, prm = case_inputs(name)
        node_M, leaf_M = oracle.ref_tree_multipoles(pos, prm, tmp)
        np.savez_compressed(os.path.join(HERE, f"tree_{name}.npz"), pos_sha256=np.array(pos_digest(pos)),
                            node_M=node_M, leaf_M=leaf_M)
        print("tree", name, node_M.shape, leaf_M.shape, flush=True)


def write_forest(tmp: str):
    counts, lengths, pos, prm = forest_inputs()
    off = np.concatenate([[0], np.cumsum(counts)])
    idx, acc, cnts = [], [], []
    for b in range(len(counts)):
        pb, p = pos[off[b]:off[b + 1]], forest_box_params(prm, lengths[b])
        if not oracle.Tree(pb, p).fits_reference_buffers():
            raise SystemExit(f"forest box {b}: exceeds the reference's fixed buffers")
        a, cnt = oracle.ref_force(pb, p, tmp)
        oa, _ = oracle.force(pb, p)
        agree = float(np.max(np.linalg.norm(oa - a, axis=1) / np.linalg.norm(a, axis=1)))
        if agree > 1e-12:
            raise SystemExit(f"forest box {b}: the oracle and the reference differ by {agree:.2e}")
        k = np.arange(0, len(pb), 8)
        idx.append(off[b] + k)
        acc.append(a[k])
        cnts.append([cnt[x] for x in ("idxP2P", "idxM2L", "walk_m2l_count", "p2p_count_remote", "nleaf", "nnode")])
        print("forest box", b, len(pb), cnt, f"oracle vs reference {agree:.1e}", flush=True)
    np.savez_compressed(os.path.join(HERE, "forest.npz"), pos_sha256=np.array(pos_digest(pos)),
                        counts=counts, lengths=lengths, idx=np.concatenate(idx).astype(np.int64),
                        acc=np.concatenate(acc), counters=np.array(cnts, np.int64))


def main():
    """no arguments: every fixture; `tree` or `forest`: that fixture; otherwise only the
    named force cases"""
    if not oracle.ref_available():
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    names = sys.argv[1:] or list(FORCE_CASES)
    with tempfile.TemporaryDirectory() as tmp:
        if not sys.argv[1:]:
            write_ops(tmp)
        if sys.argv[1:] in (["tree"], ["forest"]):
            (write_tree_cases if sys.argv[1] == "tree" else write_forest)(tmp)
            return
        if not sys.argv[1:]:
            write_tree_cases(tmp)
            write_forest(tmp)
        write_force_cases(tmp, names)


if __name__ == "__main__":
    main()
