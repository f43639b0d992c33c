"""Write the interaction-list totals of the 256^3 headline workload, from the oracle.

The workload (ic.workload_clustered(nside=256, order=4): 256^3 clustered particles,
theta 0.4, leaf 32, rc = 6 rs, 26 periodic images) is regenerated from its seed; the
oracle (oracle/refcpu.c, pinned against the reference's own golden vectors at the sizes
the reference runs here) builds the whole tree and walks it without storing the lists,
counting the P2P and M2L entries (src/fmm.c:406-712 and src/remotes.c:213-552 for the
home box and the images).  About a minute of one core.

    python tests/golden/make_counts256.py

writes lists_clustered256.json next to this script: the tree's leaf and node counts, the
two list totals and a SHA-256 of the positions (to catch generator drift).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "photons-2.0_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ic  # noqa: E402
import oracle  # noqa: E402

OUT = os.path.join(HERE, "lists_clustered256.json")


def main():
    pos, prm = ic.workload_clustered(nside=256, order=4)
    t = oracle.Tree(pos, prm)
    L = oracle.lib()
    rec = {
        "workload": "ic.workload_clustered(nside=256, order=4)",
        "positions_sha256": hashlib.sha256(np.ascontiguousarray(pos).tobytes()).hexdigest(),
        "nleaf": t.nleaf,
        "nnode": t.nnode,
        "p2p_entries": int(L.orc_p2p_list(t.h, None, None, None, 0)),
        "m2l_entries": int(L.orc_m2l_list(t.h, None, None, None, 0)),
    }
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1)
        f.write("\n")
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
