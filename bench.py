"""Benchmark: short-range force evaluations of 256^3 clustered particles on one MI355X.

One step = one short-range force evaluation (tree build, P2M/M2M, dual-tree walk,
M2L, L2L, P2P + L2P) of all particles, positions already resident in HBM.  Both
precisions run on the same particles: FP32-mixed (the headline `value`) and FP64.
Per-phase GPU times come from HIP events the library records on its own stream.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--nside 256] [--order 4]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_precision(torch, pos_dev, n, prm, precision, steps, warmup):
    ctx = pg.Context(prm, precision)
    ctx.set_timing(True)
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


def cpu_baseline(nside=40):
    """The FP64 CPU oracle (a restatement of the reference force path, one core) on a
    clustered sample with the 256^3 workload's density, rs and rc (nside^3 particles)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    pos, prm = ic.workload_clustered(nside=nside, order=4)
    t0 = time.perf_counter()
    oracle.force(pos, prm)
    dt = time.perf_counter() - t0
    return {"value": len(pos) / dt, "unit": "particles/s", "cores": 1, "kind": "port",
            "sample": f"{nside}^3 clustered particles at the same density, rs, rc, theta, p=4 and leaf size "
                      f"as the 256^3 workload (box {prm.boxsize:.0f} kpc/h), one force evaluation, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nside", type=int, default=256)
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp32", "fp64", "both"], default="both")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="first copy MiB mebibytes device to device (a plain copy kernel of known size, "
                         "to scale hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each box or domain runs on one GPU);"
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
                     "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "note": "27 flops per interaction x interactions / mean P2P kernel time (HIP events)"},
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
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, one core)")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
