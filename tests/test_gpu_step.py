"""One TreePM step on device-resident particles (photons_step_device) against the
oracle's pass of the reference's time loop (src/photoNs.c:158-254): the FMM part
from oracle/refcpu.c, the PM part from oracle/pm.py, the kick and drift factors of
src/initial.c:650-683."""
import numpy as np
import pytest

import ic
import oracle
import pm
import photons_gpu as pg

pytestmark = pytest.mark.gpu

LOGA = np.log([0.10, 0.104, 0.108])


def factors():
    return [(0.5 * pg.kick_factor(li, lf) * ic.GRAV_CONST, pg.drift_factor(li, lf))
            for li, lf in zip(LOGA[:-1], LOGA[1:])]


@pytest.fixture(scope="module")
def state():
    nside = 24
    pos, prm = ic.workload_clustered(nside=nside, order=4)
    rng = np.random.Generator(np.random.PCG64(21))
    vel = rng.normal(size=pos.shape) * 300.0
    pm_force = lambda p: pm.pm_force(p, prm.mass, nside, prm.boxsize, prm.rs)
    acc0, _ = oracle.force(pos, prm)
    return dict(pos=pos, vel=vel, acc=acc0, acc_pm=pm_force(pos), prm=prm, nside=nside, pm_force=pm_force)


def run_gpu(st, precision, with_pm, nsteps):
    import torch
    prm = st["prm"]
    dev = {k: torch.from_numpy(np.ascontiguousarray(st[k])).cuda() for k in ("pos", "vel", "acc", "acc_pm")}
    torch.cuda.synchronize()
    ctx = pg.Context(prm, precision)
    mesh = pg.PM(st["nside"], prm.boxsize, prm.rs, prm.mass) if with_pm else None
    n = len(st["pos"])
    for dkh, dd in factors()[:nsteps]:
        print(f"step: n={n} precision={precision} pm={with_pm}", flush=True)
        ctx.step_device(mesh, n, dev["pos"].data_ptr(), dev["vel"].data_ptr(), dev["acc"].data_ptr(),
                        dev["acc_pm"].data_ptr() if with_pm else 0, dkh, dd)
    ctx.sync()
    out = {k: v.cpu().numpy() for k, v in dev.items()}
    ctx.close()
    if mesh is not None:
        mesh.close()
    return out


def pdiff(a, b, L):
    d = np.abs(a - b)
    return np.minimum(d, L - d)


def test_two_treepm_steps_match_the_oracle(state):
    st, prm = state, state["prm"]
    p, v, a, apm = st["pos"], st["vel"], st["acc"], st["acc_pm"]
    for dkh, dd in factors():
        p, v, a, apm = oracle.treepm_step(p, v, a, apm, prm, st["pm_force"], dkh, dd)
    got = run_gpu(st, pg.FP64, True, 2)
    L = prm.boxsize
    moved = pdiff(p, st["pos"], L).max()
    print(f"largest displacement over two steps: {moved / L:.2e} L")
    assert moved > 1e-4 * L  # the steps move particles measurably
    assert pdiff(got["pos"], p, L).max() <= 1e-10 * L
    assert np.abs(got["vel"] - v).max() <= 1e-10 * np.abs(v).max()
    e = np.linalg.norm(got["acc"] - a, axis=1) / np.linalg.norm(a, axis=1)
    assert e.max() <= 1e-10
    assert np.abs(got["acc_pm"] - apm).max() <= 1e-10 * np.abs(apm).max()


def test_without_pm_is_the_short_range_step(state):
    st, prm = state, state["prm"]
    (dkh, dd), = factors()[:1]
    p, v, a = oracle.step(st["pos"], st["vel"], st["acc"], prm, dkh, dd)
    got = run_gpu(st, pg.FP64, False, 1)
    assert pdiff(got["pos"], p, prm.boxsize).max() <= 1e-10 * prm.boxsize
    assert np.abs(got["vel"] - v).max() <= 1e-10 * np.abs(v).max()
    assert np.array_equal(got["acc_pm"], st["acc_pm"])  # untouched


def test_fp32_mixed_step_stays_close_to_fp64(state):
    a64 = run_gpu(state, pg.FP64, True, 2)
    a32 = run_gpu(state, pg.FP32_MIXED, True, 2)
    dv = np.linalg.norm(a32["vel"] - a64["vel"], axis=1) / np.linalg.norm(a64["vel"] - state["vel"], axis=1)
    print(f"FP32-mixed vs FP64 velocity change after two steps: p99 {np.percentile(dv, 99):.2e}")
    assert np.percentile(dv, 99) <= 1e-4


def test_step_refuses_bad_calls(state):
    import torch
    prm = state["prm"]
    ctx = pg.Context(prm, pg.FP64)
    mesh = pg.PM(state["nside"], prm.boxsize, prm.rs, prm.mass)
    d = torch.zeros((prm.maxleaf, 3), dtype=torch.float64, device="cuda")
    p = d.data_ptr()
    # periodic force needs more than maxleaf particles
    with pytest.raises(pg.PhotonsError, match="maxleaf"):
        ctx.step_device(None, prm.maxleaf, p, p, p, 0, 1.0, 1.0)
    big = torch.zeros((4 * prm.maxleaf, 3), dtype=torch.float64, device="cuda")
    q = big.data_ptr()
    # a PM handle needs the acc_pm array
    with pytest.raises(pg.PhotonsError, match="acc_pm"):
        ctx.step_device(mesh, 4 * prm.maxleaf, q, q, q, 0, 1.0, 1.0)
    # n = 0 is a no-op
    ctx.step_device(mesh, 0, q, q, q, q, 1.0, 1.0)
    ctx.close()
    mesh.close()
