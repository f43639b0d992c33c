"""Benchmark: short-range force evaluations of 256^3 clustered particles on one MI355X.

One step = one short-range force evaluation (tree build, P2M/M2M, dual-tree walk,
M2L, L2L, P2P + L2P) of all particles, positions already resident in HBM.  Both
precisions run on the same particles: FP32-mixed (the headline `value`) and FP64.
Per-phase GPU times come from HIP events the library records on its own stream.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--nside 256] [--order 4]

Four more workloads print their own JSON line (not the default):
    --workload ensemble   64 independent clustered boxes of 32^3..128^3 particles (log-uniform,
                          seed 5) in one batched evaluation, leaf size and P2P staging autotuned
    --workload serving    256 resident boxes of 16^3..192^3 particles (seed 6), single-step
                          TreePM requests (short-range and PM forces + kick/drift; --serve-no-pm:
                          short-range only) with Poisson arrivals at swept rates up to saturation,
                          HBM budget 200 GB: throughput and p50/p99 step latency per rate, and the
                          FP32-mixed against FP64 error of a served step
    --workload treepm     whole TreePM steps (kick, drift, short-range and PM forces, kick) of the
                          --nside^3 clustered box, particles resident in HBM
    --workload small      latency of one small clustered box (--small-sides, default 16^3 and
                          32^3): ordinary evaluations against the replayed hipGraph
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "photons-2.0_amd"))

import ic  # noqa: E402
import photons_gpu as pg  # noqa: E402

METRIC = "particles/s per force eval, FP32-mixed vs FP64 @1e-4 err; p50/p99 latency"
PEAK_FP32_VALU = 157.3  # TFLOP/s, MI355X vector FP32
PEAK_FP64_VALU = 78.6
P2P_FLOPS = 27  # per particle-particle interaction
# HBM bytes per FP32 P2P launch (k_p2p_f32<4, 256>) at 256^3 clustered, p = 4, from
# rocprofv3 --pmc FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from a 1 GiB copy in the
# same run); the counts and commands are kept with the P2P counter summary under profiles/
P2P_FP32_TRAFFIC_256 = (2.0 * 13064333 + 793660) * 1024


def log(msg):
    print(msg, file=sys.stderr, flush=True)


PLAN_REUSE = True  # --no-plan-reuse: every evaluation sizes its buffers as it goes


def run_precision(torch, pos_dev, n, prm, precision, steps, warmup):
    ctx = pg.Context(prm, precision)
    ctx.set_timing(True)
    ctx.set_plan_reuse(PLAN_REUSE)
    acc_dev = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    name = "fp32-mixed" if precision == pg.FP32_MIXED else "fp64"
    for w in range(warmup):
        ctx.force_device(n, pos_dev.data_ptr(), acc_dev.data_ptr())
        ctx.sync()
        log(f"[{name}] warmup {w + 1}/{warmup} phases {ctx.phase_ms()}")
    wall, phases = [], []
    for s in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.force_device(n, pos_dev.data_ptr(), acc_dev.data_ptr())
        ctx.sync()
        torch.cuda.synchronize()
        wall.append(time.perf_counter() - t0)
        phases.append(ctx.phase_ms())
        log(f"[{name}] step {s + 1}/{steps} {wall[-1] * 1e3:.1f} ms")
    st = ctx.stats()
    acc = acc_dev.cpu().numpy()
    ctx.close()
    ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    return dict(wall=np.array(wall), phases=ph, stats=st, acc=acc)


def _cpu_rank(args):
    """one 'rank': the FP64 oracle on its own clustered box (separate process)"""
    nside, seed = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    pos, prm = ic.workload_clustered(nside=nside, order=4, seed=seed)
    t0 = time.perf_counter()
    oracle.force(pos, prm)
    return len(pos), time.perf_counter() - t0


REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "refharness")


def _ref_rank(args):
    """one 'rank': the reference's own force path (its fmm.c, operator.c, remotes.c,
    toptree.c and utility.c compiled from its sources into oracle/_ref/refharness, one MPI
    rank) on its own clustered box, timed around the process"""
    import subprocess
    import tempfile
    nside, seed = args
    pos, prm = ic.workload_clustered(nside=nside, order=3, seed=seed)
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([len(pos), prm.maxleaf, 0, 0], np.int32).tobytes())
            f.write(np.array([prm.boxsize, prm.mass, prm.soft, prm.rs, prm.rc, prm.theta]).tobytes())
            f.write(np.ascontiguousarray(pos, np.float64).tobytes())
        t0 = time.perf_counter()
        subprocess.run([REF_HARNESS, "force", fin, fout], check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        return len(pos), time.perf_counter() - t0


def _headline_leaves(args):
    """one core: the FP64 oracle (p = 4) on the 256^3 headline box, building the whole tree
    and every multipole and evaluating only k seeded target leaves (walks, M2L, L2L, P2P,
    L2P of those leaves), timed around the call"""
    k, seed, nleaf = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    pos, prm = ic.workload_clustered(nside=256, order=4)
    # the oracle's tree is the GPU's (same leaves, tested): leaf ids below its leaf count
    sel = np.sort(np.random.default_rng(seed).choice(nleaf, k, replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    oracle.force_leaves(pos, prm, sel)
    return k, time.perf_counter() - t0


def cpu_headline(nleaf, cores):
    """The CPU side of the headline workload itself: `cores` processes at once on the GPU
    box's host, each the FP64 oracle (the reference's algorithm, p = 4, single threaded)
    on the 256^3 box, half of them evaluating 24 sampled target leaves and half 96; the two
    call times give the per-leaf cost and the tree + multipole cost, and the whole
    evaluation on one core is modelled as tree + nleaf x per-leaf, spread over `cores`
    ranks as the reference would (one rank per core, each its share of the leaves and of
    the tree)."""
    import multiprocessing as mp
    ks = [24 if i % 2 == 0 else 96 for i in range(cores)]
    ctx = mp.get_context("spawn")
    with ctx.Pool(cores) as pool:
        t0 = time.perf_counter()
        res = pool.map(_headline_leaves, [(k, 700 + i, nleaf) for i, k in enumerate(ks)])
        wall = time.perf_counter() - t0
    lo = np.mean([t for k, t in res if k == 24])
    hi = np.mean([t for k, t in res if k == 96])
    per_leaf = max((hi - lo) / 72.0, 1e-9)
    t_tree = max(lo - 24 * per_leaf, 0.0)
    t_one = t_tree + nleaf * per_leaf
    n = 256 ** 3
    return {"value": float(n / (t_one / cores)), "unit": "particles/s", "cores": cores, "kind": "port",
            "sample": f"{cores} concurrent single-threaded processes of the FP64 oracle (p=4) on the 256^3 "
                      f"headline box itself, each building the whole tree and multipoles and evaluating 24 or 96 "
                      f"seeded target leaves ({wall:.1f} s wall); per leaf {per_leaf * 1e3:.1f} ms, tree and "
                      f"multipoles {t_tree:.1f} s; the whole evaluation modelled as tree + {nleaf} leaves x per "
                      f"leaf = {t_one:.0f} s on one core, over {cores} ranks",
            "one_core": {"value": float(n / t_one), "seconds_modelled": float(t_one)}}


def lattice32_one_core():
    """The small CPU-runnable case on one core: 32^3 lattice + Gaussian displacements,
    one FP64 force evaluation (the oracle, p = 4), timed"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    pos, prm = ic.workload_lattice32(order=4)
    t0 = time.perf_counter()
    oracle.force(pos, prm)
    t = time.perf_counter() - t0
    return {"value": len(pos) / t, "unit": "particles/s", "cores": 1, "kind": "port", "seconds": t,
            "sample": "32^3 lattice + Gaussian displacements (sigma 0.2 cell), periodic, p=4, theta 0.4, leaf 32, "
                      "rc 6 rs: one FP64 force evaluation of the oracle on one core"}


def cpu_baseline(nside=48, cores=None):
    """The reference CPU force path beside the GPU: the reference's own FMM sources compiled
    here (oracle/_ref/refharness, -O2, single threaded like one reference rank; its expansion
    order is the compiled p = 3) when present, else the FP64 oracle (their restatement, p = 4;
    the two take the same time), on clustered boxes with the 256^3 workload's density, rs, rc,
    theta and leaf size (nside^3 particles each): first one rank alone, then one rank per host
    core at once, each on its own box (the reference's one rank per core, without its domain
    exchange)."""
    import multiprocessing as mp
    if cores is None:
        cores = min(16, len(os.sched_getaffinity(0)))
    reference = os.access(REF_HARNESS, os.X_OK)
    if reference:
        try:
            n1, t1 = _ref_rank((nside, 101))
        except Exception as e:  # e.g. a runtime library missing on this host
            log(f"reference harness unusable here ({e}); timing the oracle instead")
            reference = False
    rank = _ref_rank if reference else _cpu_rank
    if not reference:
        n1, t1 = rank((nside, 101))
    ctx = mp.get_context("spawn")
    with ctx.Pool(cores) as pool:
        t0 = time.perf_counter()
        res = pool.map(rank, [(nside, 101 + k) for k in range(cores)])
        wall = time.perf_counter() - t0
    ntot = sum(r[0] for r in res)
    _, prm = ic.workload_clustered(nside=nside, order=4)
    what = ("the reference's force path (its own sources, p=3 as compiled)" if reference else
            "the FP64 oracle (p=4)")
    return {"value": ntot / wall, "unit": "particles/s", "cores": cores, "kind": "reference" if reference else "port",
            "sample": f"{cores} concurrent single-threaded ranks of {what}, each one force evaluation of its own "
                      f"{nside}^3 clustered box at the density, rs, rc, theta and leaf size of the 256^3 workload "
                      f"(box {prm.boxsize:.0f} kpc/h), {wall:.1f} s wall",
            "one_core": {"value": n1 / t1, "seconds": t1}}


def run_ensemble(args, torch):
    """64 clustered boxes (32^3..128^3, log-uniform, seed 5) as one batched evaluation."""
    log("generating the ensemble")
    counts, lengths, pos, prm = ic.workload_ensemble(64, 32 ** 3, 128 ** 3, seed=5, order=args.order)
    n = int(counts.sum())
    d_pos = torch.from_numpy(np.ascontiguousarray(pos)).to("cuda")
    d_acc = torch.empty_like(d_pos)
    res = {}
    for name, prec in (("fp32", pg.FP32_MIXED), ("fp64", pg.FP64)):
        ctx = pg.Context(prm, prec)
        leaf, stage, table = ctx.autotune_boxes(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr(), reps=1)
        log(f"[{name}] autotuned leaf {leaf} staging {stage}: {table}")
        wall = []
        for s in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.force_boxes_device(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr())
            ctx.sync()
            torch.cuda.synchronize()
            if s >= args.warmup:
                wall.append(time.perf_counter() - t0)
                log(f"[{name}] step {s - args.warmup + 1}/{args.steps} {wall[-1] * 1e3:.1f} ms")
        res[name] = dict(wall=np.array(wall), acc=d_acc.cpu().numpy(), leaf=leaf, stage=stage,
                         table={f"leaf{k[0]}_stage{k[1]}": v for k, v in table.items()}, stats=ctx.stats())
        ctx.close()
    # the same evaluation planned and captured in a hipGraph (autotuned FP32 parameters):
    # replays timed on the host clock per launch + wait, and on the GPU (HIP events)
    p32 = ic.ForceParams(**{**prm.as_dict(), "maxleaf": res["fp32"]["leaf"]})
    gctx = pg.Context(p32, pg.FP32_MIXED)
    gctx.set_p2p_stage(res["fp32"]["stage"])
    gctx.set_timing(True)
    gctx.graph_create(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr(), margin=0.25)
    gwall, gms = [], []
    for s in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gctx.graph_launch()
        ovf, ms = gctx.graph_check()
        if ovf:
            raise RuntimeError(f"graph replay overflowed its plan ({ovf})")
        if s >= args.warmup:
            gwall.append(time.perf_counter() - t0)
            gms.append(ms)
            log(f"[fp32 graph] replay {s - args.warmup + 1}/{args.steps} {gwall[-1] * 1e3:.1f} ms (GPU {ms:.1f} ms)")
    graph_same = bool(np.array_equal(d_acc.cpu().numpy(), res["fp32"]["acc"]))
    gctx.graph_destroy()
    gctx.close()
    small = small_box_latency(torch, args)
    # the precision error at equal tree parameters: FP64 with the FP32 leaf size
    p64 = ic.ForceParams(**{**prm.as_dict(), "maxleaf": res["fp32"]["leaf"]})
    ref = pg.Context(p64, pg.FP64)
    ref.force_boxes_device(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr())
    ref.sync()
    a64 = d_acc.cpu().numpy()
    ref.close()
    e = np.linalg.norm(res["fp32"]["acc"] - a64, axis=1) / np.linalg.norm(a64, axis=1)
    ms = res["fp32"]["wall"] * 1e3
    return {
        "metric": METRIC, "value": n / float(np.mean(res["fp32"]["wall"])), "unit": "particles/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": float(np.mean(ms)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32-mixed",
        "data": "synthetic",
        "config": {"workload": "ensemble of 64 periodic clustered boxes, 32^3..128^3 particles (log-uniform, "
                               "seed 5), density of 256^3 in 500 Mpc/h, one batched evaluation, autotuned",
                   "particles": n, "boxes": 64, "order": prm.order, "theta": prm.theta,
                   "leaf_fp32": res["fp32"]["leaf"], "p2p_staging_fp32": res["fp32"]["stage"],
                   "leaf_fp64": res["fp64"]["leaf"]},
        "latency_ms": {"p50": float(np.percentile(ms, 50)), "p99": float(np.percentile(ms, 99))},
        "autotune_ms": {"fp32": res["fp32"]["table"], "fp64": res["fp64"]["table"]},
        "fp64": {"value": n / float(np.mean(res["fp64"]["wall"])),
                 "ms_per_step": float(np.mean(res["fp64"]["wall"]) * 1e3)},
        "fp32_over_fp64": float(np.mean(res["fp64"]["wall"]) / np.mean(res["fp32"]["wall"])),
        "error_fp32_vs_fp64": {"p50": float(np.median(e)), "p99": float(np.percentile(e, 99)), "max": float(e.max())},
        "p2p_interactions": res["fp32"]["stats"]["p2p_interactions"],
        "graph": {"what": "the FP32-mixed evaluation of all 64 boxes planned (capacities 25% above one measured "
                          "evaluation) and captured in one hipGraph; one graph launch per evaluation, no host "
                          "round trip inside it",
                  "ms_per_step": float(np.mean(gwall) * 1e3), "gpu_ms": float(np.mean(gms)),
                  "value": n / float(np.mean(gwall)), "same_as_ordinary": graph_same},
        "small_box_latency_ms": small,
    }


def small_box_latency(torch, args):
    """one clustered box of 16^3 / 32^3 particles (--small-sides): ordinary evaluations
    (planned from the previous one, a wait at the end of each) against the replayed
    graph, wall clock per evaluation"""
    out = {}
    for ns in args.small_sides:
        n = ns ** 3
        box = 500000.0 * (n / 256.0 ** 3) ** (1.0 / 3.0)
        pos, _ = ic.clustered(n, box, seed=1)
        prm = ic.force_params(n, box, ns, rc_factor=6.0, order=args.order)
        d_pos = torch.from_numpy(np.ascontiguousarray(pos)).to("cuda")
        d_acc = torch.empty_like(d_pos)
        counts, lengths = np.array([n]), np.array([box])
        ctx = pg.Context(prm, pg.FP32_MIXED)
        reps = 50
        for k in range(reps + 5):
            if k == 5:
                ctx.sync()
                t0 = time.perf_counter()
            ctx.force_boxes_device(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr())
        ctx.sync()
        t_ord = (time.perf_counter() - t0) / reps * 1e3
        ctx.graph_create(counts, lengths, d_pos.data_ptr(), d_acc.data_ptr(), margin=0.25)
        for k in range(reps + 5):
            if k == 5:
                ctx.graph_check()
                t0 = time.perf_counter()
            ctx.graph_launch()
        ctx.graph_check()
        t_graph = (time.perf_counter() - t0) / reps * 1e3
        ctx.close()
        out[f"{ns}^3"] = {"ordinary": t_ord, "graph": t_graph}
        log(f"{ns}^3 box: ordinary {t_ord:.3f} ms, replayed graph {t_graph:.3f} ms per evaluation")
    return out


def run_small(args, torch):
    """the small-box latencies alone (FP32-mixed)"""
    small = small_box_latency(torch, args)
    ns = args.small_sides[0]
    ms = small[f"{ns}^3"]["graph"]
    return {
        "metric": METRIC, "value": ns ** 3 / (ms * 1e-3), "unit": "particles/s", "n_gpus": 1,
        "steps": 50, "warmup": 5, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-mixed", "data": "synthetic",
        "config": {"workload": f"one clustered box of {ns}^3 particles at the density of the 256^3 workload, "
                               "planned and replayed as a hipGraph", "particles": ns ** 3, "order": args.order},
        "small_box_latency_ms": small,
    }


def run_treepm(args, torch):
    """TreePM steps (photons_step_device: both half kicks, drift, wrap, new tree and
    short-range force, PM force on an nside^3 mesh) of nside^3 clustered particles that
    stay in HBM; FP32-mixed and FP64 short-range force, FP64 PM and kick/drift."""
    log("generating particles")
    pos, prm = ic.workload_clustered(nside=args.nside, order=args.order)
    n = len(pos)
    vel = np.random.Generator(np.random.PCG64(9)).normal(size=pos.shape) * 300.0
    loga = np.log(0.5) + 0.01 * np.arange(args.warmup + args.steps + 1)
    fac = [(0.5 * pg.kick_factor(li, lf) * ic.GRAV_CONST, pg.drift_factor(li, lf)) for li, lf in zip(loga[:-1], loga[1:])]
    res = {}
    for name, prec in (("fp32", pg.FP32_MIXED), ("fp64", pg.FP64)):
        d = {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda") for k, v in (("pos", pos), ("vel", vel))}
        d["acc"] = torch.zeros_like(d["pos"])
        d["apm"] = torch.zeros_like(d["pos"])
        torch.cuda.synchronize()
        ctx = pg.Context(prm, prec)
        mesh = pg.PM(args.nside, prm.boxsize, prm.rs, prm.mass)
        ptr = {k: v.data_ptr() for k, v in d.items()}
        mesh.force_device(n, ptr["pos"], ptr["apm"])
        ctx.force_device(n, ptr["pos"], ptr["acc"])
        mesh.sync()
        ctx.sync()
        t0 = time.perf_counter()
        mesh.force_device(n, ptr["pos"], ptr["apm"])
        mesh.sync()
        pm_ms = (time.perf_counter() - t0) * 1e3
        wall = []
        for s, (dkh, dd) in enumerate(fac):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.step_device(mesh, n, ptr["pos"], ptr["vel"], ptr["acc"], ptr["apm"], dkh, dd)
            ctx.sync()
            if s >= args.warmup:
                wall.append(time.perf_counter() - t0)
                log(f"[{name}] TreePM step {s - args.warmup + 1}/{args.steps} {wall[-1] * 1e3:.1f} ms")
        res[name] = dict(wall=np.array(wall), pm_ms=pm_ms)
        ctx.close()
        mesh.close()
    ms = res["fp32"]["wall"] * 1e3
    return {
        "metric": METRIC, "value": n / float(np.mean(res["fp32"]["wall"])), "unit": "particles/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": float(np.mean(ms)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32-mixed",
        "data": "synthetic",
        "config": {"workload": f"TreePM steps of {args.nside}^3 clustered particles resident in HBM: half kicks, "
                               f"drift, wrap, new tree and short-range force, PM on a {args.nside}^3 mesh "
                               "(FP64), half kicks; a from 0.5 in steps of 0.01 in ln a",
                   "particles": n, "order": prm.order, "theta": prm.theta, "maxleaf": prm.maxleaf},
        "latency_ms": {"p50": float(np.percentile(ms, 50)), "p99": float(np.percentile(ms, 99))},
        "pm_ms": res["fp32"]["pm_ms"],
        "fp64": {"value": n / float(np.mean(res["fp64"]["wall"])),
                 "ms_per_step": float(np.mean(res["fp64"]["wall"]) * 1e3)},
    }


MESH_CLASSES = (16, 32, 48, 64, 96, 128, 192, 256)


def serving_mesh(L, rs):
    """a served box's PM mesh: the reference's NSIDE = 1.25 L / rs rounded up to one of a
    few sizes (a lane keeps one PM context per mesh size)"""
    want = 1.25 * L / rs
    return next((m for m in MESH_CLASSES if m >= want - 1e-9), MESH_CLASSES[-1])


def serving_step_error(prm, boxes, dkh, dd, pm_on):
    """The FP32-mixed error of a served step: the same boxes, one step each, in an
    FP32-mixed server; the short-range force each box holds afterwards against the FP64
    force at the positions it holds (an FP64 context on those positions: the same tree,
    so only the precision differs)"""
    srv = pg.Server(prm, pg.FP32_MIXED, nlanes=2, hbm_budget=200e9, max_batch_particles=64 << 20)
    ids = [srv.add_box(p, v, L, nside=serving_mesh(L, prm.rs) if pm_on else None) for p, v, L in boxes]
    srv.drain()
    for r in [srv.submit(b, dkh, dd) for b in ids]:
        srv.wait(r)
    got = [srv.read_box(b) for b in ids]
    srv.close()
    e = []
    for (p32, v32, a32), (_, _, L) in zip(got, boxes):
        want = pg.Context(ic.ForceParams(**{**prm.as_dict(), "boxsize": float(L)}), pg.FP64).force(p32)
        e.append(np.linalg.norm(a32 - want, axis=1) / np.linalg.norm(want, axis=1))
    e = np.concatenate(e)
    return {"boxes": len(boxes), "particles": int(len(e)),
            "force": {"p50": float(np.median(e)), "p99": float(np.percentile(e, 99)), "max": float(e.max())}}


def run_serving(args, torch):
    """Resident boxes, Poisson arrivals of single-step requests at swept rates (each
    precision's sweep stops once the server falls behind the offered rate)."""
    nbox = args.serve_boxes
    pm_on = not args.serve_no_pm
    log(f"generating {nbox} boxes")
    counts, lengths, pos, prm = ic.workload_ensemble(nbox, 16 ** 3, 192 ** 3, seed=6, order=args.order)
    off = np.concatenate([[0], np.cumsum(counts)])
    vel = ic.normal(7, pos.size).reshape(pos.shape) * 300.0
    dkh = 0.5 * pg.kick_factor(np.log(0.30), np.log(0.31)) * ic.GRAV_CONST
    dd = pg.drift_factor(np.log(0.30), np.log(0.31))
    out = {}
    for name, prec, rates in (("fp32", pg.FP32_MIXED, args.rates32), ("fp64", pg.FP64, args.rates64)):
        srv = pg.Server(prm, prec, nlanes=2, hbm_budget=200e9, max_batch_particles=64 << 20)
        t0 = time.perf_counter()
        ids = [srv.add_box(pos[off[b]:off[b + 1]], vel[off[b]:off[b + 1]], lengths[b],
                           nside=serving_mesh(lengths[b], prm.rs) if pm_on else None) for b in range(nbox)]
        srv.drain()  # the first force of every box
        log(f"[{name}] {nbox} boxes resident ({int(counts.sum())} particles), first forces in "
            f"{time.perf_counter() - t0:.1f} s")
        rows = []
        rng = np.random.Generator(np.random.PCG64(8))
        for rate in rates:
            nreq = max(32, int(rate * args.serve_seconds))
            boxes = rng.integers(0, nbox, nreq)
            gaps = rng.exponential(1.0 / rate, nreq)
            lat, done = [None] * nreq, threading.Event()
            rids = [None] * nreq
            submitted = threading.Semaphore(0)

            def waiter():
                for k in range(nreq):
                    submitted.acquire()
                    lat[k] = srv.wait(rids[k])
                done.set()

            th = threading.Thread(target=waiter)
            th.start()
            s0 = srv.stats()
            t_start = time.perf_counter()
            t_next = t_start
            for k in range(nreq):
                t_next += gaps[k]
                dt = t_next - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                rids[k] = srv.submit(ids[boxes[k]], dkh, dd)
                submitted.release()
            done.wait()
            th.join()
            span = time.perf_counter() - t_start
            s1 = srv.stats()
            stepped = s1["particles_stepped"] - s0["particles_stepped"]
            lat_a = np.array(lat)
            row = {"rate_req_per_s": rate, "requests": nreq, "particles_per_s": stepped / span,
                   "achieved_req_per_s": nreq / span, "latency_ms": {"p50": float(np.percentile(lat_a, 50)),
                                                                     "p99": float(np.percentile(lat_a, 99))},
                   "mean_boxes_per_batch": ((s1["mean_boxes_per_batch"] * s1["batches"] -
                                             s0["mean_boxes_per_batch"] * s0["batches"]) /
                                            max(1, s1["batches"] - s0["batches"]))}
            row["max_lanes_busy"] = s1["max_lanes_busy"]
            rows.append(row)
            log(f"[{name}] rate {rate}/s: {row['particles_per_s'] / 1e6:.1f} M particles/s, "
                f"{row['achieved_req_per_s']:.1f} req/s, "
                f"latency p50 {row['latency_ms']['p50']:.0f} ms p99 {row['latency_ms']['p99']:.0f} ms, "
                f"{row['mean_boxes_per_batch']:.1f} boxes per batch")
            if row["achieved_req_per_s"] < 0.8 * rate:
                break  # saturated: the server falls behind the offered rate
        out[name] = rows
        srv.close()
    log("serving: FP32-mixed against FP64 steps of the first 16 boxes")
    err = serving_step_error(prm, [(pos[off[b]:off[b + 1]], vel[off[b]:off[b + 1]], lengths[b])
                                   for b in range(min(16, nbox))], dkh, dd, pm_on)
    best32 = max(r["particles_per_s"] for r in out["fp32"])
    best64 = max(r["particles_per_s"] for r in out["fp64"])
    sat = max(out["fp32"], key=lambda r: r["particles_per_s"])
    return {
        "metric": METRIC, "value": best32, "unit": "particles/s", "n_gpus": 1, "steps": None, "warmup": None,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-mixed", "data": "synthetic",
        "config": {"workload": f"serving: {nbox} resident clustered boxes of 16^3..192^3 particles (log-uniform, "
                               "seed 6), Poisson single-step requests (short-range force"
                               + (", PM force on a mesh of about 1.25 L / rs cells per side" if pm_on else "")
                               + " + FP64 kick/drift), 2 lanes, HBM budget 200 GB",
                   "particles": int(counts.sum()), "boxes": nbox, "pm": pm_on},
        "latency_ms": sat["latency_ms"], "sweep": out, "fp64": {"value": best64},
        "fp32_over_fp64": best32 / best64, "error_fp32_vs_fp64_per_step": err,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nside", type=int, default=256)
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp32", "fp64", "both"], default="both")
    ap.add_argument("--workload", choices=["clustered", "ensemble", "serving", "treepm", "small"],
                    default="clustered")
    ap.add_argument("--small-sides", type=int, nargs="+", default=[16, 32],
                    help="particles per side of the small boxes (latency)")
    ap.add_argument("--serve-boxes", type=int, default=256)
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="arrival span per rate")
    ap.add_argument("--rates32", type=float, nargs="+", default=[4, 8, 16, 32, 64, 128, 256, 512])
    ap.add_argument("--rates64", type=float, nargs="+", default=[2, 4, 8, 16, 32, 64])
    ap.add_argument("--serve-no-pm", action="store_true", help="served steps without the PM part")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="first copy MiB mebibytes device to device (a plain copy kernel of known size, "
                         "to scale hardware byte counters)")
    ap.add_argument("--no-plan-reuse", action="store_true",
                    help="size every evaluation's buffers as it goes (host reads per tree level and walk depth)")
    args = ap.parse_args()
    global PLAN_REUSE
    PLAN_REUSE = not args.no_plan_reuse
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (each box or domain runs on one GPU);"
              " run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {nbytes} bytes read and written once")
        del src, dst
    if args.workload != "clustered":
        fn = {"ensemble": run_ensemble, "serving": run_serving, "treepm": run_treepm,
              "small": run_small}[args.workload]
        print(json.dumps(fn(args, torch)), flush=True)
        return
    log("generating particles")
    pos, prm = ic.workload_clustered(nside=args.nside, order=args.order)
    n = len(pos)
    pos_dev = torch.from_numpy(np.ascontiguousarray(pos)).to("cuda")
    torch.cuda.synchronize()
    res = {}
    if args.only in ("fp32", "both"):
        res["fp32"] = run_precision(torch, pos_dev, n, prm, pg.FP32_MIXED, args.steps, args.warmup)
    if args.only in ("fp64", "both"):
        res["fp64"] = run_precision(torch, pos_dev, n, prm, pg.FP64, args.steps, args.warmup)
    head = res["fp32"] if "fp32" in res else res["fp64"]
    ms = head["wall"] * 1e3
    st = head["stats"]
    p2p_ms = head["phases"]["p2p"]
    inter = st["p2p_interactions"]
    fp32 = "fp32" in res
    achieved = P2P_FLOPS * inter / (p2p_ms * 1e-3) / 1e12
    peak = PEAK_FP32_VALU if fp32 else PEAK_FP64_VALU
    out = {
        "metric": METRIC,
        "value": n / float(np.mean(head["wall"])),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(ms)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed" if fp32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.nside}^3 clustered particles (50% in Plummer spheres, dN/dM ~ M^-1.9), "
                               f"periodic {prm.boxsize / 1000:.0f} Mpc/h box, PM grid {args.nside}^3, rs 1.25 cells, "
                               f"rc 6 rs, theta {prm.theta}, p={prm.order}, leaf {prm.maxleaf}",
                   "particles": n, "order": prm.order, "theta": prm.theta, "maxleaf": prm.maxleaf,
                   "rc_over_rs": prm.rc / prm.rs},
        "latency_ms": {"p50": float(np.percentile(ms, 50)), "p99": float(np.percentile(ms, 99))},
        "phases_ms": head["phases"],
        "interactions_per_s": inter / (p2p_ms * 1e-3),
        "p2p_interactions": inter,
        "tree": {"nleaf": st["nleaf"], "nnode": st["nnode"], "depth": st["depth"],
                 "p2p_pairs": st["p2p_pairs"], "m2l_pairs": st["m2l_pairs"]},
        "roofline": {"bound": "valu", "kernel": "k_p2p_f32" if fp32 else "k_p2p_f64", "achieved": achieved,
                     "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": (P2P_FP32_TRAFFIC_256 if fp32 and args.nside == 256 and prm.order == 4 else None),
                     "note": "27 flops per interaction x interactions / mean P2P kernel time (HIP events on the "
                             "library's stream); compute bound on the vector ALU: ~24 VALU instructions per "
                             "interaction in the pair loop, one of them transcendental (rsqrt; the split factor "
                             "is a cubic read from an LDS table); packed FP32 is not used, so 78.6 of the "
                             "157.3 TFLOP/s peak is reachable at most; traffic = HBM bytes per launch from "
                             "rocprofv3 PMC"},
    }
    if "fp32" in res and "fp64" in res:
        a32, a64 = res["fp32"]["acc"], res["fp64"]["acc"]
        e = np.linalg.norm(a32 - a64, axis=1) / np.linalg.norm(a64, axis=1)
        out["fp64"] = {"value": n / float(np.mean(res["fp64"]["wall"])),
                       "ms_per_step": float(np.mean(res["fp64"]["wall"]) * 1e3),
                       "latency_ms": {"p50": float(np.percentile(res["fp64"]["wall"] * 1e3, 50)),
                                      "p99": float(np.percentile(res["fp64"]["wall"] * 1e3, 99))},
                       "phases_ms": res["fp64"]["phases"]}
        out["fp32_over_fp64"] = out["value"] / out["fp64"]["value"]
        out["error_fp32_vs_fp64"] = {"p50": float(np.median(e)), "p99": float(np.percentile(e, 99)),
                                     "max": float(e.max())}
    if not args.no_cpu_baseline and args.nside == 256:
        cores = min(16, len(os.sched_getaffinity(0)))
        log(f"cpu baseline: the oracle on the 256^3 box itself ({cores} processes)")
        cb = cpu_headline(int(st["nleaf"]), cores)
        log("cpu baseline: the reference's own force path on small boxes, and the 32^3 lattice on one core")
        cb["reference_small_boxes"] = cpu_baseline(cores=cores)
        cb["lattice32_one_core"] = lattice32_one_core()
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
