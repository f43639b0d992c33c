"""Named inputs of the golden force vectors (shared by make_golden.py and the tests)."""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "photons-2.0_amd"))

import ic  # noqa: E402


def _clustered(nside, nside_pm, rc_factor=6.0, seed=1, theta=0.4, maxleaf=32):
    n = nside ** 3
    box = 500000.0 * (n / 256.0 ** 3) ** (1.0 / 3.0)
    pos, _ = ic.clustered(n, box, seed=seed)
    return pos, ic.force_params(n, box, nside_pm, rc_factor=rc_factor, theta=theta, maxleaf=maxleaf)


def _lattice(rc_factor, maxleaf=32, theta=0.4):
    pos, prm = ic.workload_lattice32(rc_factor=rc_factor)
    prm.maxleaf = maxleaf
    prm.theta = theta
    return pos, prm


def _lcdm():
    pos = np.load(os.path.join(HERE, "lcdm_pos.npy")).astype(np.float64)
    return pos, ic.force_params(len(pos), 100000.0, 64, rc_factor=6.0)


FORCE_CASES = {
    # 32^3 lattice + 0.2-cell Gaussian, 100 Mpc/h, PM 64^3, theta 0.4, leaf 32
    "lattice32_rc4.5": lambda: _lattice(4.5),
    "lattice32_rc6": lambda: _lattice(6.0),
    "lattice32_leaf8_th0.5": lambda: _lattice(6.0, maxleaf=8, theta=0.5),
    # clustered (half in Plummer spheres), density of 256^3 in 500 Mpc/h
    "clus16_pm32": lambda: _clustered(16, 32),
    "clus24_pm48": lambda: _clustered(24, 48),
    "clus32_pm32": lambda: _clustered(32, 32),
    "clus24_pm48_th0.3": lambda: _clustered(24, 48, theta=0.3, seed=7),
    # the reference's own demo initial conditions (demo/ic_lcdm.gdt2)
    "lcdm_demo": _lcdm,
    # 262,144 clustered particles: the parity case at scale (every 64th particle stored)
    "clus64_pm64": lambda: _clustered(64, 64),
    # other leaf size, opening angle and cutoff on clustered inputs
    "clus32_leaf16_th0.5_rc4.5": lambda: _clustered(32, 32, rc_factor=4.5, seed=11, theta=0.5, maxleaf=16),
    "clus48_pm48_th0.3": lambda: _clustered(48, 48, seed=13, theta=0.3),
}

# stored stride of the sampled accelerations (default 8)
STRIDE = {"clus64_pm64": 64, "clus48_pm48_th0.3": 32}

# cases whose oracle run takes about a minute on one core: the CPU suite checks them
# only with --full (the reference and the oracle agreed to 5e-15 on them when the
# vectors were written); the GPU parity test compares with the stored vectors
LARGE_CASES = {"clus64_pm64", "clus48_pm48_th0.3"}


def case_inputs(name):
    return FORCE_CASES[name]()


# The forest case: three clustered boxes of 16^3..24^3 particles (log-uniform, seeded) at
# the density of the 256^3 workload, evaluated together; one explicit parameter set for
# every box (mass, softening, rs and rc of 256^3 particles in 500 Mpc/h, p = 3), each
# box with its own length.  The golden vectors hold each box's reference force alone.
def forest_inputs():
    counts, lengths, pos, prm = ic.workload_ensemble(3, 16 ** 3, 24 ** 3, seed=31, order=3)
    return counts, lengths, pos, prm


def forest_box_params(prm, length):
    return ic.ForceParams(**{**prm.as_dict(), "boxsize": float(length)})


# the multipoles of every cell after fmm_prepare, for walk_m2m
TREE_CASES = ("clus16_pm32",)
