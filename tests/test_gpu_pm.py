"""The PM long-range force on the GPU (photons_pm_*) against the numpy restatement
of the reference's partmesh()/convolution (oracle/pm.py, itself pinned by closed
forms in test_pm_oracle.py)."""
import numpy as np
import pytest

import ic
import pm
import photons_gpu as pg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nside,kind", [(32, "random"), (32, "clustered"), (64, "lattice")])
def test_matches_the_oracle(nside, kind):
    L = 100000.0
    rs = 1.25 * L / nside
    if kind == "random":
        pos = np.random.Generator(np.random.PCG64(3)).random((20000, 3)) * L
    elif kind == "clustered":
        pos, _ = ic.clustered(20000, L, seed=4)
    else:
        pos = ic.lattice_gauss(32, L, 0.2, seed=0)
    mass = 211.75
    want = pm.pm_force(pos, mass, nside, L, rs)
    got = pg.PM(nside, L, rs, mass).force(pos)
    scale = np.abs(want).max()
    err = np.abs(got - want).max() / scale
    print(f"{kind} N={nside}: max |a - a_oracle| / max |a| = {err:.2e}")
    assert err <= 1e-11
    assert np.abs(got.sum(axis=0)).max() <= 1e-10 * scale


def test_full_size_mesh_and_device_path():
    import time
    import torch
    pos, prm = ic.workload_clustered(nside=256, order=4)
    d_pos = torch.from_numpy(np.ascontiguousarray(pos)).cuda()
    d_acc = torch.zeros_like(d_pos)
    p = pg.PM(256, prm.boxsize, prm.rs, prm.mass)
    p.force_device(len(pos), d_pos.data_ptr(), d_acc.data_ptr())
    p.sync()
    t0 = time.perf_counter()
    for _ in range(3):
        p.force_device(len(pos), d_pos.data_ptr(), d_acc.data_ptr())
    p.sync()
    ms = (time.perf_counter() - t0) / 3 * 1e3
    acc = d_acc.cpu().numpy()
    print(f"PM 256^3 mesh, 256^3 clustered particles: {ms:.1f} ms per evaluation")
    assert np.isfinite(acc).all()
    assert np.abs(acc.sum(axis=0)).max() <= 1e-9 * np.abs(acc).max()
    # the host path gives the same answer up to the order of the deposit's atomic adds
    sub = slice(0, 100000)
    host = p.force(pos)
    assert np.abs(host[sub] - acc[sub]).max() <= 1e-10 * np.abs(acc).max()
