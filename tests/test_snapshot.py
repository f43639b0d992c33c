"""Gadget-2 snapshot/IC files (host code of the library; no GPU needed).

Pinned by the reference's own fixture demo/ic_lcdm.gdt2 (32^3 particles, L = 100
Mpc/h, m = 211.7538, z = 49) through the positions committed in
tests/golden/lcdm_pos.npy (read there by an independent numpy parser), by round trips,
and by the a^1.5 velocity convention of src/snapshot.c:255, 472.
"""
import os
import struct

import numpy as np
import pytest

import photons_gpu as pg

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_IC = "/root/reference/demo/ic_lcdm.gdt2"


@pytest.mark.skipif(not os.path.exists(REF_IC), reason="the reference's demo files are not on this machine")
def test_reads_the_reference_demo_ic():
    h = pg.snapshot_header(REF_IC)
    assert h["n"] == 32768 and h["npart"][1] == 32768 and h["npart_total"][1] == 32768
    assert h["boxsize"] == 100000.0
    assert abs(h["mass"][1] - 211.7538) < 1e-3
    assert abs(h["redshift"] - 49.0) < 1e-9
    pos, vel = pg.snapshot_read(REF_IC)
    want = np.load(os.path.join(GOLDEN, "lcdm_pos.npy"), allow_pickle=False)
    assert np.array_equal(pos, want.astype(np.float64))
    assert (pos >= 0).all() and (pos < h["boxsize"]).all()
    # a sub-range reads the same particles
    p2, v2 = pg.snapshot_read(REF_IC, 1000, 500)
    assert np.array_equal(p2, pos[1000:1500]) and np.array_equal(v2, vel[1000:1500])


def test_round_trip_and_layout(tmp_path):
    rng = np.random.Generator(np.random.PCG64(5))
    n, L, z = 1000, 5.0e4, 9.0
    pos = rng.random((n, 3)) * L
    vel = rng.normal(size=(n, 3)) * 200.0
    path = tmp_path / "snap.gdt2"
    pg.snapshot_write(path, pos, vel, L, z, 12.5, omega_m=0.3, omega_l=0.7, hubble=0.68)
    raw = open(path, "rb").read()
    assert len(raw) == (4 + 256 + 4) + 2 * (4 + 12 * n + 4)
    assert struct.unpack_from("<i", raw, 0)[0] == 256 and struct.unpack_from("<i", raw, 264)[0] == 12 * n
    h = pg.snapshot_header(path)
    assert h["n"] == n and h["boxsize"] == L and h["redshift"] == z and h["mass"][1] == 12.5
    assert abs(h["time"] - 0.1) < 1e-15 and h["omega_m"] == 0.3 and h["hubble"] == 0.68
    p, v = pg.snapshot_read(path)
    assert np.array_equal(p, pos.astype(np.float32).astype(np.float64))
    # velocities are stored as v / a^1.5 in float32 and scaled back on reading
    a15 = (1.0 / (1.0 + z)) ** 1.5
    stored = np.frombuffer(raw, dtype="<f4", count=3 * n, offset=264 + 8 + 12 * n + 4).reshape(n, 3)
    assert np.allclose(stored, vel / a15, rtol=1e-6)
    assert np.allclose(v, vel, rtol=2e-7, atol=0)


def test_errors(tmp_path):
    with pytest.raises(pg.PhotonsError, match="cannot open"):
        pg.snapshot_header(tmp_path / "missing")
    path = tmp_path / "s.gdt2"
    pg.snapshot_write(path, np.zeros((10, 3)), np.zeros((10, 3)), 1.0, 0.0, 1.0)
    with pytest.raises(pg.PhotonsError, match="outside"):
        pg.snapshot_read(path, 5, 10)
    open(tmp_path / "short", "wb").write(b"\x00" * 100)
    with pytest.raises(pg.PhotonsError, match="short"):
        pg.snapshot_header(tmp_path / "short")
