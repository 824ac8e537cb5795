"""Autotuner of the batched replicas: size classes and wavefronts per replica (WPR).

The batched kernel runs one persistent launch per size class; each workgroup of a class
holds WPR wavefronts and owns one replica at a time.  Small replicas want one wavefront
(many resident workgroups per CU, little reduction cost); large replicas want more
wavefronts per workgroup (their candidate sets and LDS footprints are larger, and fewer
workgroups fit on a CU anyway).  The best split depends on the size distribution, so the
tuner measures it:

  1. for every candidate class boundary list, each class is timed alone for every WPR
     in {1, 2, 4, 8} on the replicas of the workload that fall into it, and the fastest
     WPR is kept;
  2. the full workload is then timed with every boundary list and its tuned WPRs, and the
     list with the highest events/s wins.

The result goes to jellyfysh_amd/tuning/gfx950.json, which DeviceBatch reads by default.

    python -m jellyfysh_amd.autotune [--replicas 4096] [--legs 1000] [--out PATH]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from .batch import TUNING_FILE, DeviceBatch, batched_lj_workload

WPRS = (1, 2, 4, 8)
SPLITS = (
    (16384,),
    (256, 16384),
    (512, 16384),
    (256, 768, 16384),
    (192, 512, 1024, 16384),
    (128, 256, 512, 1024, 16384),
)


def _rate(batch, legs, reps):
    batch.run_graph(legs, 1)                          # warm-up (and graph instantiation)
    s0 = batch.states()["legs"].sum()
    ms = batch.run_graph(legs, reps)
    s1 = batch.states()["legs"].sum()
    return float(s1 - s0) / (ms * 1e-3), ms


def _subset(n, lengths, pos, sel):
    return n[sel], lengths[sel], [pos[i] for i in np.flatnonzero(sel)]


def tune(n, lengths, pos, legs=1000, reps=2, seed=17, log=print):
    n = np.asarray(n)
    lengths = np.asarray(lengths)
    best_wpr = {}
    per_class = []
    for split in SPLITS:
        lo = 0
        for hi in split:
            key = (lo, hi)
            sel = (n > lo) & (n <= hi)
            if key not in best_wpr and sel.any():
                sn, sl, sp = _subset(n, lengths, pos, sel)
                rates = {}
                for w in WPRS:
                    b = DeviceBatch(sn, sl, sp, seed=seed, classes=([hi], [w]))
                    rates[w] = _rate(b, legs, reps)[0]
                    b.close()
                best_wpr[key] = max(rates, key=rates.get)
                per_class.append(dict(lo=lo, hi=hi, replicas=int(sel.sum()),
                                      events_per_sec={str(k): v for k, v in rates.items()}))
                log("class (%d, %d]: %d replicas, events/s by WPR %s -> %d"
                    % (lo, hi, int(sel.sum()), {k: "%.3g" % v for k, v in rates.items()}, best_wpr[key]))
            lo = hi
    results = []
    for split in SPLITS:
        lo, wprs = 0, []
        for hi in split:
            wprs.append(best_wpr.get((lo, hi), 1))
            lo = hi
        b = DeviceBatch(n, lengths, pos, seed=seed, classes=(list(split), wprs))
        rate, ms = _rate(b, legs, reps)
        b.close()
        results.append(dict(class_max_n=list(split), class_wpr=wprs, events_per_sec=rate))
        log("split %s WPR %s: %.4g events/s" % (list(split), wprs, rate))
    best = max(results, key=lambda r: r["events_per_sec"])
    return best, results, per_class


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--replicas", type=int, default=4096)
    ap.add_argument("--legs", type=int, default=1000)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=TUNING_FILE)
    args = ap.parse_args(argv)
    import torch
    n, lengths, pos = batched_lj_workload(args.replicas, seed=2024)
    t0 = time.time()
    best, results, per_class = tune(n, lengths, pos, args.legs, args.reps,
                                    log=lambda m: print("[autotune] " + m, flush=True))
    out = dict(device=torch.cuda.get_device_name(0), workload="batched LJ 12-6, %d replicas, N 64-2048 "
               "log-uniform, rho 0.7-0.9, %d legs per timing" % (args.replicas, args.legs),
               class_max_n=best["class_max_n"], class_wpr=best["class_wpr"], events_per_sec=best["events_per_sec"],
               splits=results, classes=per_class, seconds=time.time() - t0)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(dict(class_max_n=best["class_max_n"], class_wpr=best["class_wpr"],
                          events_per_sec=best["events_per_sec"])), flush=True)


if __name__ == "__main__":
    sys.exit(main())
