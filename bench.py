#!/usr/bin/env python3
"""ECMC events/sec on one MI355X: batched replicas vs one plain chain, with pressures.

Workload of the headline number (one "step" = every replica advances --legs legs):
  4096 independent LJ 12-6 replicas (two inverse-power factors, truncated and shifted at
  rc = 2.5 sigma, eps = sigma = 1, beta = 1/1.5), N drawn log-uniformly in [64, 2048],
  rho uniformly in [0.7, 0.9], chains of 1.0 sigma cycling +x,+y,+z; synthetic lattice
  starts.  Inputs are resident on the GPU before the timed region.
Beside it, the plain form of the same potential at N = 4096, rho = 0.8 (FCC start) runs
the reference's leg loop on one workgroup, and the oracle runs a bounded sample of the
batched workload on the host cores (the CPU baseline).

    python bench.py [--steps K] [--warmup W] [--legs L]
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ECMC events/sec on 1 MI355X, batched vs plain, at matched pressure (2σ)"
FLOPS_PER_INVERSION = 60.0     # FP64-equivalent operations of one inverse-power inversion
PEAK_TFLOPS = 157.3            # MI355X dense FP32 peak


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--legs", type=int, default=2000, help="legs per replica per step")
    ap.add_argument("--replicas", type=int, default=4096)
    ap.add_argument("--plain-legs", type=int, default=20000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    return ap.parse_args()


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def plain_chain_rate(legs):
    """Plain form: one chain of N = 4096 LJ at rho = 0.8 with 6^3 cells, on one workgroup."""
    from jellyfysh_amd.chain import DeviceChain, make_desc
    n_cells = 16
    n = 4 * n_cells ** 3
    L = (n / 0.8) ** (1 / 3)
    rng = np.random.default_rng(7)
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    a = L / n_cells
    pts = np.array([(np.array([i, j, k]) + b) * a for i in range(n_cells) for j in range(n_cells)
                    for k in range(n_cells) for b in base])
    pos = np.mod(pts + rng.uniform(-0.01, 0.01, pts.shape), L)
    pc = ["lj_twelve", "lj_six", "cell_boundary"]
    taggers = [dict(tag="lj_twelve", kind="pair_cells", potential="inverse_power_truncated", params=[12.0, 4.0, 2.5],
                    create=pc, trash=pc),
               dict(tag="lj_six", kind="pair_cells", potential="inverse_power_truncated", params=[6.0, -4.0, 2.5],
                    create=pc, trash=pc),
               dict(tag="cell_boundary", kind="cell_boundary", create=pc, trash=pc),
               dict(tag="end_of_chain", kind="end_of_chain", create=["end_of_chain"] + pc,
                    trash=["end_of_chain"] + pc, chain_time=1.0),
               dict(tag="end_of_run", kind="end_of_run", create=["end_of_run"],
                    trash=["end_of_chain"] + pc + ["start_of_run", "end_of_run"], end_of_run_time=1e12),
               dict(tag="start_of_run", kind="start_of_run", create=pc + ["end_of_chain", "end_of_run"],
                    trash=["start_of_run"])]
    desc = make_desc(3, n, L, 1 / 1.5, int(L / 2.5), taggers, seed=5, scheduler="heap")
    chain = DeviceChain(desc, pos)
    chain.run(2000, log=False)                      # warm-up
    e0 = chain.estimators()
    t0 = time.perf_counter()
    done = chain.run(legs, log=False)
    wall = time.perf_counter() - t0
    ms = chain.last_kernel_ms()
    est = chain.estimators() - e0
    return dict(events=int(done), kernel_ms=ms, wall_s=wall, events_per_sec=done / (ms * 1e-3),
                pressure=1.0 + est[0] / est[1] if est[1] > 0 else None, n=n)


def cpu_baseline(n, lengths, pos, seconds, threads):
    """The oracle's batched restatement (FP64, glibc math, every candidate inverted) on host
    threads, one replica per thread at a time, for about `seconds` of wall time."""
    from oracle import oracle as O
    O.lib()
    threads = threads or min(16, os.cpu_count() or 1)
    order = np.argsort(-np.asarray(n))[:: max(1, len(n) // (4 * threads))][: 4 * threads]
    # calibrate legs so that the sample takes about `seconds`
    probe = O.BatchedReplica(pos[order[0]], float(lengths[order[0]]), replica=int(order[0]), seed=1,
                             math_mode=O.MATH_LIBM)
    t = time.perf_counter()
    probe.run(200, log=False)
    per_leg = (time.perf_counter() - t) / 200
    legs = max(200, int(seconds * threads / (len(order) * per_leg * 0.7)))

    def work(r):
        rep = O.BatchedReplica(pos[r], float(lengths[r]), replica=int(r), seed=1, math_mode=O.MATH_LIBM)
        return rep.run(legs, log=False)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        done = sum(ex.map(work, [int(r) for r in order]))
    wall = time.perf_counter() - t0
    return dict(value=done / wall, unit="events/s", cores=threads, kind="port",
                sample="%d replicas of the batched workload (largest-first stride sample, N %d-%d) x %d legs, "
                       "oracle restatement in FP64 with glibc math, %d host threads, %.1f s"
                       % (len(order), int(np.min(np.asarray(n)[order])), int(np.max(np.asarray(n)[order])), legs,
                          threads, wall))


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: each replica runs independently, and a multi-GPU run would "
              "split the replica list across devices with no exchange; --gpus %d is not run here." % args.gpus)
        sys.exit(2)
    import torch
    from jellyfysh_amd import _lib
    from jellyfysh_amd.batch import DeviceBatch, batched_lj_workload, default_classes, pressure
    _lib.require_device()
    n, lengths, pos = batched_lj_workload(args.replicas, seed=2024)
    log("workload: %d replicas, %d units" % (len(n), int(np.sum(n))))
    batch = DeviceBatch(n, lengths, pos, seed=17, classes=default_classes())
    for _ in range(args.warmup):
        batch.run_graph(args.legs, 1)
    s0 = batch.states()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms = batch.run_graph(args.legs, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s1 = batch.states()
    events = int((s1["legs"] - s0["legs"]).sum())
    value = events / wall
    log("batched: %d events in %.3f s (kernel %.3f ms)" % (events, wall, kernel_ms))
    # algorithmic work: every factor candidate of every leg inverted once
    cands = float((s1["stats"][:, 0] - s0["stats"][:, 0]).sum())
    n_factors = 2
    flops = cands * n_factors * FLOPS_PER_INVERSION
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    exact = float((s1["stats"][:, 3] - s0["stats"][:, 3]).sum())
    est = s1["est"] - s0["est"]
    p = pressure(est)
    rho = n / lengths ** 3
    sel = np.abs(rho - 0.8) < 0.01
    out = {
        "metric": METRIC, "value": value, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched LJ 12-6 (rc 2.5, shifted): %d replicas, N log-uniform 64-2048, rho 0.7-0.9, "
                               "beta 1/1.5, chain 1.0, %d legs per replica per step" % (len(n), args.legs),
                   "replicas": int(len(n)), "units": int(np.sum(n)), "legs_per_step": args.legs,
                   "classes": {"max_n": [int(v) for v in batch.classes[0]],
                               "wavefronts_per_replica": [int(v) for v in batch.classes[1]]}},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS, "traffic": None,
                     "note": "compute-bound VALU kernel priced against the FP32 peak; algorithmic work = %g "
                             "FLOP per factor-candidate inversion (%d factors), exact FP64 inversions per "
                             "candidate = %.4f" % (FLOPS_PER_INVERSION, n_factors, exact / max(cands, 1))},
        "batched": {"kernel_ms": kernel_ms, "events": events, "candidates_per_leg": cands / max(events, 1),
                    "pressure_rho_0.80": {"mean": float(np.mean(p[sel])) if sel.any() else None,
                                          "err": float(np.std(p[sel], ddof=1) / np.sqrt(sel.sum()))
                                          if sel.sum() > 1 else None, "replicas": int(sel.sum())}},
    }
    if not args.no_plain:
        log("plain chain ...")
        pl = plain_chain_rate(args.plain_legs)
        out["plain"] = pl
        out["batched_over_plain"] = value / pl["events_per_sec"]
    if not args.no_cpu:
        log("cpu baseline ...")
        cb = cpu_baseline(n, lengths, pos, args.cpu_seconds, args.cpu_threads)
        out["cpu_baseline"] = cb
        out["batched_over_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
