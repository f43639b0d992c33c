"""The reference program's run on one GPU (examples/photons_run): a Gadget-2 IC and a
parameter file in the reference's format, TreePM steps, a Gadget-2 snapshot out.
Checked against the oracle's steps (oracle.treepm_step: refcpu FMM + oracle/pm.py PM)
from the same IC as the program reads it; the output file holds float32, so the
comparison is to float32 rounding."""
import math
import os
import subprocess

import numpy as np
import pytest

import ic
import oracle
import pm
import photons_gpu as pg

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "photons-2.0_amd", "examples", "photons_run")


def test_run_matches_the_oracle(tmp_path):
    nside, z0, af, nstep = 16, 9.0, 0.108, 2
    pos, prm = ic.workload_clustered(nside=nside, order=3, theta=0.4)
    L, n = prm.boxsize, len(pos)
    vel = np.random.Generator(np.random.PCG64(31)).normal(size=pos.shape) * 300.0
    ic_path = tmp_path / "ic.gdt2"
    pg.snapshot_write(ic_path, pos, vel, boxsize=L, redshift=z0, mass=prm.mass)
    (tmp_path / "run.param").write_text(
        f"# test run\nInputPath {ic_path}\nSnapFormat 2\nSnapNumber 1\nOutputPath {tmp_path}\n"
        f"OutputName run\nOPENANGLE 0.4\nNumMeshSide {nside}\nGravConstant {ic.GRAV_CONST}\n")
    print(f"photons_run: n={n} nside={nside} nstep={nstep}", flush=True)
    r = subprocess.run([EXE, str(tmp_path / "run.param"), "--nstep", str(nstep), "--af", str(af)],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0

    # the oracle from the IC as the program reads it (float32 in the file)
    p, v = pg.snapshot_read(ic_path)
    h = pg.snapshot_header(ic_path)
    rs = 1.25 * L / nside
    oprm = ic.ForceParams(boxsize=L, mass=h["mass"][1], soft=0.03 * L / math.pow(n, 0.3333333), rs=rs,
                          rc=4.5 * rs, theta=0.4, maxleaf=32, order=3, periodic=1)
    pm_force = lambda x: pm.pm_force(x, oprm.mass, nside, L, rs)
    a, _ = oracle.force(p, oprm)
    apm = pm_force(p)
    ai = 1.0 / (1.0 + z0)
    dl = (math.log(af) - math.log(ai)) / nstep
    for k in range(nstep):
        li, lf = math.log(ai) + k * dl, math.log(ai) + (k + 1) * dl
        dkh = 0.5 * oracle.kick_loga(li, lf, h["omega_m"], h["omega_l"]) * ic.GRAV_CONST
        p, v, a, apm = oracle.treepm_step(p, v, a, apm, oprm, pm_force, dkh, oracle.drift_loga(li, lf, h["omega_m"],
                                                                                               h["omega_l"]))
    out = tmp_path / "run_3.0"
    ho = pg.snapshot_header(out)
    assert ho["n"] == n and abs(ho["redshift"] - (1 / af - 1)) < 1e-12 and ho["boxsize"] == L
    gp, gv = pg.snapshot_read(out)
    d = np.abs(gp - p)
    d = np.minimum(d, L - d)
    moved = np.abs(p - pos)
    print(f"largest displacement {np.minimum(moved, L - moved).max() / L:.2e} L; "
          f"program vs oracle: pos {d.max() / L:.2e} L, vel {np.abs(gv - v).max() / np.abs(v).max():.2e}")
    assert d.max() <= 2e-7 * L
    assert np.abs(gv - v).max() <= 1e-6 * np.abs(v).max()
    assert np.abs(gv - vel).max() > 1e-2 * np.abs(vel).max()  # the kicks changed the velocities


def test_run_options_fp32_and_no_pm(tmp_path):
    """--fp32 and --no-pm: the FP32-mixed run stays within the FP32 tolerance of the FP64
    run; without PM the result is the short-range-only steps of the oracle"""
    nside, z0, af, nstep = 16, 9.0, 0.104, 1
    pos, prm = ic.workload_clustered(nside=nside, order=3, theta=0.4)
    L, n = prm.boxsize, len(pos)
    vel = np.random.Generator(np.random.PCG64(32)).normal(size=pos.shape) * 300.0
    ic_path = tmp_path / "ic.gdt2"
    pg.snapshot_write(ic_path, pos, vel, boxsize=L, redshift=z0, mass=prm.mass)
    outs = {}
    for tag, extra in (("fp64", []), ("fp32", ["--fp32"]), ("nopm", ["--no-pm"])):
        (tmp_path / f"{tag}.param").write_text(
            f"InputPath {ic_path}\nOutputPath {tmp_path}\nOutputName {tag}\nOPENANGLE 0.4\nNumMeshSide {nside}\n"
            f"GravConstant {ic.GRAV_CONST}\nSOFTENING 40.0\nSPLITSCALE {1.25 * L / nside}\n")
        print(f"photons_run {tag}", flush=True)
        r = subprocess.run([EXE, str(tmp_path / f"{tag}.param"), "--nstep", str(nstep), "--af", str(af)] + extra,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert "soft = 40" in r.stdout
        outs[tag] = pg.snapshot_read(tmp_path / f"{tag}_3.0")
    dv = np.linalg.norm(outs["fp32"][1] - outs["fp64"][1], axis=1)
    assert np.percentile(dv, 99) <= 1e-4 * np.abs(outs["fp64"][1]).max()
    # the short-range-only step against the oracle's (float32 file rounding)
    p, v = pg.snapshot_read(ic_path)
    h = pg.snapshot_header(ic_path)
    rs = 1.25 * L / nside
    oprm = ic.ForceParams(boxsize=L, mass=h["mass"][1], soft=40.0, rs=rs, rc=4.5 * rs, theta=0.4, maxleaf=32,
                          order=3, periodic=1)
    a, _ = oracle.force(p, oprm)
    li, lf = math.log(1 / (1 + z0)), math.log(af)
    dkh = 0.5 * oracle.kick_loga(li, lf, h["omega_m"], h["omega_l"]) * ic.GRAV_CONST
    p, v, a = oracle.step(p, v, a, oprm, dkh, oracle.drift_loga(li, lf, h["omega_m"], h["omega_l"]))
    d = np.abs(outs["nopm"][0] - p)
    assert np.minimum(d, L - d).max() <= 2e-7 * L
    assert np.abs(outs["nopm"][1] - v).max() <= 1e-6 * np.abs(v).max()
