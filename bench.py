#!/usr/bin/env python3
"""ECMC events/sec on one MI355X: batched replicas vs one plain chain, with pressures.

Workload of the headline number (one "step" = every replica advances --legs legs):
  4096 independent LJ 12-6 replicas (two inverse-power factors, truncated and shifted at
  rc = 2.5 sigma, eps = sigma = 1, beta = 1/1.5), N drawn log-uniformly in [64, 2048],
  rho uniformly in [0.7, 0.9], chains of 1.0 sigma cycling +x,+y,+z; synthetic lattice
  starts.  Inputs are resident on the GPU before the timed region.
Beside it, the plain form of the same potential at N = 4096, rho = 0.8 (16^3 lattice start) runs
the reference's leg loop on one workgroup, and the oracle runs a bounded sample of the
batched workload on the host cores (the CPU baseline).  256 charged-soft-sphere replicas
(N = 512, merged-image Coulomb + inverse power 9) report their rate in one launch of the plain
chain kernel against one replica alone, once with the far cells through the cell-bounding
potential and once through the cell veto (--no-charged skips them).

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
PEAK_TFLOPS = 157.3            # MI355X FP32 vector peak
PEAK_FP64_TFLOPS = 78.6       # MI355X FP64 vector peak
PEAK_HBM_GBS = 8000.0         # MI355X HBM3E bandwidth


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--legs", type=int, default=10000, help="legs per replica per step")
    ap.add_argument("--replicas", type=int, default=4096)
    ap.add_argument("--plain-legs", type=int, default=20000, help="legs per timed chunk of the plain chain")
    ap.add_argument("--plain-seconds", type=float, default=30.0, help="plain-chain time for its pressure")
    ap.add_argument("--production-steps", type=int, default=4, help="untimed steps whose legs give the pressure")
    ap.add_argument("--plain-equilibration", type=int, default=600000,
                    help="legs of the batched kernel that equilibrate the plain chain's start")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-charged", action="store_true", help="skip the charged-soft-sphere replicas")
    ap.add_argument("--charged-replicas", type=int, default=256)
    ap.add_argument("--charged-legs", type=int, default=2000)
    ap.add_argument("--only-charged", action="store_true",
                    help="only the charged-soft-sphere lines (cell bounding, then cell veto), one JSON line")
    ap.add_argument("--only-plain", action="store_true", help="only the plain N = 4096 chain, one JSON line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel 10 times over BYTES (scales hardware byte counters)")
    return ap.parse_args()


def copy_calibration(nbytes):
    """Ten launches of a device-to-device copy of nbytes: FETCH_SIZE / WRITE_SIZE per launch
    of this kernel read nbytes each when the counters are exact."""
    import ctypes
    import torch
    from jellyfysh_amd import _lib
    _lib.require_device()
    nbytes -= nbytes % 16
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for _ in range(10):
        _lib.check(_lib.lib().ecmc_copy_bytes(_lib.dptr(src), _lib.dptr(dst), nbytes, stream))
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
    print(json.dumps({"copy_kernel_bytes_per_launch": nbytes, "launches": 10}), flush=True)


def profiled_traffic(kernel, legs, profile="batched_kernel_lj_4096_replicas_counters.txt", tag=None):
    """HBM bytes per launch of `kernel` at `legs` legs per replica (and workload `tag`), from the
    committed counter profile of this workload (FETCH_SIZE scaled by the copy calibration's
    factor, plus WRITE_SIZE); None when the profile has no line for it."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        # traffic <kernel> [tag=<T>] legs=<L> bytes_per_launch=<B>
        if not f or f[0] != "traffic" or len(f) < 4 or f[1] != kernel:
            continue
        fields = dict(x.split("=", 1) for x in f[2:] if "=" in x)
        if fields.get("legs") == str(legs) and fields.get("tag") == tag:
            return float(fields["bytes_per_launch"])
    return None


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def plain_chain_rate(legs, pressure_seconds, equilibration_legs):
    """Plain form: one chain of N = 4096 LJ at rho = 0.8 with 6^3 cells, on one workgroup.

    The start is equilibrated first: the 16^3 lattice (N = 4096 is not 4 m^3, so no FCC
    lattice fits; spacing L / 16 = 1.07722 sigma, jitter +-0.01 sigma) is run through the
    batched kernel as a single replica of the same system for `equilibration_legs` legs,
    then handed to the plain chain."""
    from jellyfysh_amd.batch import DeviceBatch, lattice_positions
    from jellyfysh_amd.chain import DeviceChain, make_desc
    n = 4096
    L = (n / 0.8) ** (1 / 3)
    pos = lattice_positions(n, L, 0.01, np.random.default_rng(7))
    eq = DeviceBatch([n], [L], [pos], seed=23, classes=([16384], [4]))
    done_eq, t_eq = 0, time.perf_counter()
    while done_eq < equilibration_legs:
        eq.run(min(100000, equilibration_legs - done_eq))
        done_eq = int(eq.states()["legs"][0])
        log("plain start: %d equilibration legs" % done_eq)
    pos = eq.positions()[0]
    eq.close()
    t_eq = time.perf_counter() - t_eq
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
    warm = chain.run(2000)                          # warm-up, logged: candidates per leg
    ncand = float(warm["ncand"].mean())
    # the timed rate: one chunk of `legs` legs
    t0 = time.perf_counter()
    done = chain.run(legs, log=False)
    wall = time.perf_counter() - t0
    ms = chain.last_kernel_ms()
    out = dict(n=n, rho=0.8, start="16^3 grid (spacing 1.07722, jitter 0.01) equilibrated by %d legs of the batched "
                                   "kernel (%.1f s)" % (done_eq, t_eq),
               events=int(done), kernel_ms=ms, wall_s=wall, events_per_sec=done / (ms * 1e-3))
    lds_b, lds_p, lds_r = chain.launch_lds()
    out["launch_lds"] = dict(bytes=lds_b, positions=lds_p, ring=lds_r)
    kname = "ecmc::chain_kernel<3,%d>" % chain.kernel_variant()[1]   # the specialized leg loop that ran
    out["roofline"] = plain_roofline(kname, ncand, 3, 27, done, ms, charged=False)
    out["roofline"]["traffic"] = profiled_traffic(kname, legs, "chain_kernel_plain_lj_4096_counters.txt", "plain")
    out["cpu_single_core"] = plain_cpu_single_core(desc, chain.positions(), seconds=4.0)
    # pressure: chunks of `legs` legs for `pressure_seconds`; the standard error comes from
    # 8 blocks of consecutive chunks (chunks of one chain are correlated)
    chunk_z, t_end = [], time.perf_counter() + pressure_seconds
    while time.perf_counter() < t_end:
        e1 = chain.estimators()
        chain.run(legs, log=False)
        d = chain.estimators() - e1
        if d[1] > 0:
            chunk_z.append(1.0 + d[0] / d[1])
        log("plain chunk %d: Z = %.4f" % (len(chunk_z), chunk_z[-1] if chunk_z else float("nan")))
    z = np.asarray(chunk_z)
    nb = min(8, len(z))
    blocks = np.array([b.mean() for b in np.array_split(z, nb)]) if nb else np.zeros(0)
    out["pressure"] = dict(mean=float(z.mean()) if len(z) else None,
                           err=float(blocks.std(ddof=1) / np.sqrt(nb)) if nb > 1 else None,
                           chunks=int(len(z)), legs_per_chunk=int(legs), blocks=int(nb))
    return out


def plain_roofline(kernel, ncand, dim, cells, events, ms, charged):
    """Algorithmic HBM bytes of the plain per-chain form: every candidate's position (dim x 8 B),
    identifier (4 B) and, with charges, charge (8 B), the nearby-cell table (8 B per cell), the
    active unit's write-back (dim x 8 B) and the 32 B event record, per committed leg; achieved =
    that x legs / kernel time of the launch, against the HBM peak.  The kernel is bound by the
    latency of one workgroup per chain, not by bandwidth, so the fraction is small."""
    per_leg = ncand * (dim * 8 + 4 + (8 if charged else 0)) + cells * 8 + dim * 8 + 32
    achieved = per_leg * events / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
            "kernel": kernel, "bytes_per_leg": per_leg, "candidates_per_leg": ncand, "legs_per_launch": int(events),
            "launch_ms": ms, "traffic": None,
            "note": "one workgroup per chain; latency-bound (per-leg dependent steps between barriers)"}


def plain_cpu_single_core(desc, positions, seconds, charges=None):
    """The oracle's plain engine (the reference's leg loop in FP64 with glibc math, heap
    scheduler) on one host thread from the same configuration and start, for about `seconds`."""
    from oracle import oracle as O
    O.lib()
    cfg = _oracle_config_of(desc)
    e = O.Engine(cfg, positions, charges)
    e.run(200, log=False)
    legs, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        legs += e.run(500, log=False)
    wall = time.perf_counter() - t0
    return dict(value=legs / wall, unit="events/s", cores=1, kind="port",
                sample="%d legs of the same chain (same start, same wiring) on one host thread, oracle in FP64 with "
                       "glibc math" % legs)


def _oracle_config_of(desc):
    """The oracle's configuration of a device run description (same fields, same order)."""
    from oracle import oracle as O
    c = O.Config()
    c.dim, c.n, c.system_length, c.beta = desc.dim, desc.n, desc.system_length, desc.beta
    for d in range(3):
        c.cells_per_side[d] = desc.cells_per_side[d]
    c.neighbor_layers = desc.neighbor_layers
    c.max_occupants = desc.max_occupants
    c.scheduler = desc.scheduler
    c.math_mode = O.MATH_LIBM
    c.seed = desc.seed
    c.rng_skip = desc.rng_skip
    c.leaves_per_root = desc.leaves_per_root
    c.n_taggers = desc.n_taggers
    for i in range(desc.n_taggers):
        g, h = c.taggers[i], desc.taggers[i]
        for f in ("kind", "event_handler", "potential", "bounding_potential", "use_charge", "n_create", "n_trash",
                  "n_handlers", "value", "ivalue", "ivalue2", "fmap_local"):
            setattr(g, f, getattr(h, f))
        for j in range(4):
            g.par[j], g.bpar[j] = h.par[j], h.bpar[j]
        for j in range(len(h.create)):
            g.create[j], g.trash[j] = h.create[j], h.trash[j]
        for k in range(len(h.fmap_n)):
            g.fmap_n[k] = h.fmap_n[k]
            for j in range(len(h.fmap_target[k])):
                g.fmap_target[k][j] = h.fmap_target[k][j]
    return c


def batched_pressure_at(rho, z, rho0=0.8, width=0.02, n=None, n_min=256, n_max=512):
    """beta P / rho at rho0 from the replicas with |rho - rho0| < width and n_min <= N <= n_max:
    least-squares line through (rho, Z), its value at rho0 and that value's standard error."""
    sel = np.abs(rho - rho0) < width
    if n is not None:
        sel &= (n >= n_min) & (n <= n_max)
    x, y = rho[sel] - rho0, z[sel]
    if len(x) < 4:
        return dict(mean=None, err=None, replicas=int(len(x)))
    A = np.stack([np.ones_like(x), x], 1)
    coef, res, _, _ = np.linalg.lstsq(A, y, rcond=None)
    dof = len(x) - 2
    s2 = float(np.sum((y - A @ coef) ** 2)) / dof
    cov = s2 * np.linalg.inv(A.T @ A)
    return dict(mean=float(coef[0]), err=float(np.sqrt(cov[0, 0])), slope=float(coef[1]), replicas=int(len(x)),
                window="|rho - %.2f| < %.2f, %d <= N <= %d" % (rho0, width, n_min, n_max))


def charged_spheres_rate(replicas, legs, far="cell_bounding"):
    """Charged soft spheres (N = 512, L = 8, merged-image Coulomb + inverse power 9, beta = 1):
    `replicas` independent systems advanced together by the plain chain kernel (one workgroup
    each, one launch), against one replica alone.  `far`: the far cells through the
    cell-bounding potential (one candidate per occupied far cell) or the cell veto (one
    sampled cell per leg)."""
    import torch
    from jellyfysh_amd.chain import run_together
    from jellyfysh_amd.workloads import charged_soft_sphere_chains
    t0 = time.perf_counter()
    chains = charged_soft_sphere_chains(replicas, seed=1000, far=far)
    create_s = time.perf_counter() - t0
    run_together(chains, 200)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k, ms = run_together(chains, legs)
    wall = time.perf_counter() - t0
    k1, ms1 = run_together(chains[:1], legs)
    est = np.array([c.estimators() for c in chains])
    lds_b, lds_p, lds_r = chains[0].launch_lds()
    kname = "ecmc::chain_many_kernel<3,%d>" % chains[0].kernel_variant()[1]
    for c in chains:
        c.close()
    # candidates per leg from a logged run of one more replica of the same system
    from jellyfysh_amd.workloads import charged_soft_spheres, make_chain
    sysd = charged_soft_spheres(seed=999, far=far)
    probe = make_chain(sysd)
    ncand = float(probe.run(500)["ncand"].mean())
    probe.close()
    roof = plain_roofline(kname, ncand, 3, 27, int(k.sum()), ms, charged=True)
    roof["traffic"] = profiled_traffic(kname, legs, "chain_many_kernel_charged_soft_spheres_counters.txt", far)
    from jellyfysh_amd.chain import make_desc
    desc = make_desc(3, 512, sysd["system_length"], sysd["beta"], sysd["cells_per_side"], sysd["taggers"], 999,
                     scheduler="heap", max_occupants=1)
    cpu1 = plain_cpu_single_core(desc, sysd["positions"], seconds=4.0, charges=sysd["charges"])
    return dict(replicas=replicas, n=512, system_length=8.0, far_cells=far, legs_per_replica=legs,
                roofline=roof, cpu_single_core=cpu1, launch_lds=dict(bytes=lds_b, positions=lds_p, ring=lds_r),
                events_per_sec=float(k.sum()) / wall, kernel_ms=ms,
                single_replica_events_per_sec=float(k1[0]) / (ms1 * 1e-3),
                lifts_per_leg=float(est[:, 4].sum() / max(k.sum(), 1)),
                rejected_thinning_per_leg=float(est[:, 6].sum() / max(k.sum(), 1)), create_s=create_s,
                chain_displacement_per_sec=float(est[:, 1].sum()) / wall,
                note="plain chain kernel, one workgroup per replica in one launch; FP64 event times; Coulomb "
                     "far cells through %s (FP64 Ewald tables), nearby and surplus bounds with thinning; "
                     "repulsion inverted exactly over all pairs" % far.replace("_", " "))


def flexible_dipoles_rate(replicas, legs):
    """Flexible dipoles (jellyfysh_amd/configs/flexible_dipoles.ini: 500 dipoles of charges +-0.5,
    bond p = 2, r0 = 1, k = 400 within each dipole, merged-image Coulomb between dipoles with
    thinning, r^-12 repulsion between all charges of different dipoles, L = 21.544):
    `replicas` independent systems (the INI's random start with seeds 0 ..) advanced together
    by the plain chain kernel, against one alone."""
    from jellyfysh_amd.chain import DeviceChain, run_together
    from jellyfysh_amd.config import RunConfig
    cfg = RunConfig.from_file(os.path.join(ROOT, "jellyfysh_amd", "configs", "flexible_dipoles.ini"))
    t0 = time.perf_counter()
    chains, starts = [], []
    for seed in range(replicas):
        p, q, w = cfg.initial_state(seed)
        chains.append(DeviceChain(cfg.desc(len(p), seed, rng_skip=w), p, q))
        starts.append((p, q, w))
    create_s = time.perf_counter() - t0
    run_together(chains, 100)
    t0 = time.perf_counter()
    k, ms = run_together(chains, legs)
    wall = time.perf_counter() - t0
    k1, ms1 = run_together(chains[:1], legs)
    lds_b, lds_p, lds_r = chains[0].launch_lds()
    kname = "ecmc::chain_many_kernel<3,%d>" % chains[0].kernel_variant()[1]
    probe = chains[1].run(300)
    ncand = float(probe["ncand"][1:].mean()) if len(probe) > 1 else 0.0
    desc = chains[0].desc
    for c in chains:
        c.close()
    roof = plain_roofline(kname, ncand, 3, 0, int(k.sum()), ms, charged=True)
    roof["traffic"] = profiled_traffic(kname, legs, "chain_many_kernel_charged_soft_spheres_counters.txt",
                                       "flexible_dipoles")
    p, q, w = starts[0]
    cpu1 = plain_cpu_single_core(desc, p, seconds=4.0, charges=q)
    return dict(replicas=replicas, n=len(starts[0][0]), legs_per_replica=legs, roofline=roof, cpu_single_core=cpu1,
                launch_lds=dict(bytes=lds_b, positions=lds_p, ring=lds_r), events_per_sec=float(k.sum()) / wall,
                kernel_ms=ms, single_replica_events_per_sec=float(k1[0]) / (ms1 * 1e-3), create_s=create_s,
                note="plain chain kernel, one workgroup per replica in one launch; leaf units of composite "
                     "objects with factor type maps; Coulomb through the inverse-power Coulomb bound with "
                     "thinning (FP64 Ewald derivative), bond and repulsion inverted exactly (FP32 brackets "
                     "first for the repulsion)")


def cpu_baseline(n, lengths, pos, seconds, threads):
    """The oracle's batched restatement (FP64, glibc math, every candidate inverted) on host
    threads, one replica per thread at a time, for about `seconds` of wall time.  The sample
    is the 4 x threads replicas of the workload with 256 <= N <= 512 whose density is closest
    to 0.8 (their candidate sets, 255 to ~370 units, bracket the workload's mean of ~335);
    each runs its legs in two halves, and the second halves give the host's pressure."""
    from oracle import oracle as O
    O.lib()
    threads = threads or min(16, os.cpu_count() or 1)
    n = np.asarray(n)
    rho = n / np.asarray(lengths) ** 3
    pool = np.flatnonzero((n >= 256) & (n <= 512))
    order = pool[np.argsort(np.abs(rho[pool] - 0.8), kind="stable")[: 4 * threads]]
    order = order[np.argsort(-n[order], kind="stable")]          # largest first
    # one replica on one thread: the single-core rate, and the legs per replica that make the
    # sample take about `seconds`
    probe = O.BatchedReplica(pos[order[0]], float(lengths[order[0]]), replica=int(order[0]), seed=1,
                             math_mode=O.MATH_LIBM)
    probe.run(200, log=False)
    probe_legs = 3000
    t = time.perf_counter()
    probe.run(probe_legs, log=False)
    per_leg = (time.perf_counter() - t) / probe_legs
    half = max(200, int(seconds * threads / (2 * len(order) * per_leg * 1.1)))

    def work(r):
        rep = O.BatchedReplica(pos[r], float(lengths[r]), replica=int(r), seed=1, math_mode=O.MATH_LIBM)
        k = rep.run(half, log=False)
        e0 = np.array(rep.state.est[:])
        k += rep.run(half, log=False)
        return k, np.array(rep.state.est[:]) - e0

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(work, [int(r) for r in order]))
    wall = time.perf_counter() - t0
    done = sum(k for k, _ in res)
    est = np.stack([e for _, e in res])
    z = 1.0 + est[:, 0] / est[:, 1]
    pz = batched_pressure_at(rho[order], z, width=1.0, n=n[order])
    single = dict(value=1.0 / per_leg, unit="events/s", cores=1,
                  sample="one replica (N %d) x %d legs on one host thread, after 200" % (int(n[order[0]]), probe_legs))
    return dict(value=done / wall, unit="events/s", cores=threads, kind="port", single_core=single,
                sample="%d replicas of the batched workload (256 <= N <= 512, density closest to 0.8, N %d-%d) "
                       "x %d legs each, oracle restatement in FP64 with glibc math, %d host threads, %.1f s"
                       % (len(order), int(n[order].min()), int(n[order].max()), 2 * half, threads, wall),
                note="the batched restatement (Philox draws, no scheduler, no Time bookkeeping, every candidate "
                     "inverted): cheaper per leg than the reference's leg loop, so the GPU/CPU ratio is "
                     "conservative; the plain lines carry the reference-algorithm oracle timed on one core",
                pressure_rho_0_80=pz)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.only_charged:
        from jellyfysh_amd import _lib
        _lib.require_device()
        out = {}
        for far in ("cell_bounding", "cell_veto"):
            log("charged soft spheres, far cells through %s ..." % far)
            out[far] = charged_spheres_rate(args.charged_replicas, args.charged_legs, far=far)
        log("flexible dipoles ...")
        fd = flexible_dipoles_rate(64, args.charged_legs)
        print(json.dumps({"charged_soft_spheres": out, "flexible_dipoles": fd}), flush=True)
        return
    if args.only_plain:
        from jellyfysh_amd import _lib
        _lib.require_device()
        print(json.dumps({"plain": plain_chain_rate(args.plain_legs, args.plain_seconds, args.plain_equilibration)}),
              flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: each replica runs independently, and a multi-GPU run would "
              "split the replica list across devices with no exchange; --gpus %d is not run here." % args.gpus)
        sys.exit(2)
    import torch
    from jellyfysh_amd import _lib
    from jellyfysh_amd.batch import DeviceBatch, default_classes, pressure
    from jellyfysh_amd.config import BatchedRunConfig
    _lib.require_device()
    cfg = BatchedRunConfig.from_file(os.path.join(ROOT, "jellyfysh_amd", "configs", "batched_lennard_jones.ini"))
    cfg.number_of_replicas = args.replicas
    n, lengths, pos = cfg.workload()
    log("workload: %d replicas, %d units" % (len(n), int(np.sum(n))))
    batch = DeviceBatch(n, lengths, pos, classes=default_classes(), **cfg.batch_arguments())
    for i in range(args.warmup):
        batch.run_graph(args.legs, 1)
        log("warm-up step %d done" % (i + 1))
    s0 = batch.states()
    # timed steps: direct launches with events on every class's stream (a step of this size is
    # hundreds of milliseconds, so a graph replay would save nothing measurable)
    batch.set_class_timing(True)
    cls_ms = [[0.0, 0] for _ in batch.classes[0]]
    kernel_ms = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms += batch.run(args.legs)
        for c, (ms, k) in enumerate(batch.class_ms()):
            cls_ms[c][0] += ms
            cls_ms[c][1] += k
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    batch.set_class_timing(False)
    s1 = batch.states()
    events = int((s1["legs"] - s0["legs"]).sum())
    value = events / wall
    log("batched: %d events in %.3f s (steps %.3f ms)" % (events, wall, kernel_ms))
    # Roofline of the dominant kernel (the size class with the longest summed kernel time):
    # algorithmic work = every factor candidate of its replicas' legs inverted once, per launch,
    # over its average launch duration from the class's own stream events.
    cls_of = np.searchsorted(np.asarray(batch.classes[0]), n, side="left")
    dom = int(np.argmax([ms for ms, _ in cls_ms]))
    dom_ms, dom_launches = cls_ms[dom]
    cands_r = (s1["stats"][:, 0] - s0["stats"][:, 0]).astype(np.float64)
    cands = float(cands_r.sum())
    n_factors = 2
    flops_per_launch = float(cands_r[cls_of == dom].sum()) * n_factors * FLOPS_PER_INVERSION / max(dom_launches, 1)
    launch_ms = dom_ms / max(dom_launches, 1)
    achieved = flops_per_launch / (launch_ms * 1e-3) / 1e12
    exact = float((s1["stats"][:, 3] - s0["stats"][:, 3]).sum())
    dom_name = "ecmc::batched_kernel<%d>" % batch.classes[1][dom]
    traffic = profiled_traffic(dom_name, args.legs)
    # pressure from untimed production steps after the timed region (each replica then has
    # (warmup + steps) x legs legs behind it)
    batch.run_graph(args.legs, args.production_steps)
    s2 = batch.states()
    z = pressure(s2["est"] - s1["est"])
    rho = n / lengths ** 3
    pz = batched_pressure_at(rho, z, n=n)
    pz["production_legs_per_replica"] = args.legs * args.production_steps
    out = {
        "metric": METRIC, "value": value, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched LJ 12-6 (rc 2.5, shifted): %d replicas, N log-uniform 64-2048, rho 0.7-0.9, "
                               "beta 1/1.5, chain 1.0, %d legs per replica per step" % (len(n), args.legs),
                   "replicas": int(len(n)), "units": int(np.sum(n)), "legs_per_step": args.legs,
                   "classes": {"max_n": [int(v) for v in batch.classes[0]],
                               "wavefronts_per_replica": [int(v) for v in batch.classes[1]]}},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS, "traffic": traffic,
                     "peak_fp64": PEAK_FP64_TFLOPS, "frac_fp64": achieved / PEAK_FP64_TFLOPS,
                     "kernel": dom_name, "launch_ms": launch_ms, "launches": dom_launches,
                     "flop_per_launch": flops_per_launch,
                     "note": "VALU kernel (no MFMA: nothing here is a dense contraction) priced against the FP32 "
                             "vector peak, the ceiling of its FP32 filter, and against the FP64 vector peak (frac_fp64), "
                             "the ceiling of its exact inversions; algorithmic work = %g "
                             "FLOP per factor-candidate inversion (%d factors), exact FP64 inversions per "
                             "candidate = %.4f; traffic = FETCH_SIZE x 2 + WRITE_SIZE per launch from the "
                             "committed counter profile of this workload" % (FLOPS_PER_INVERSION, n_factors,
                                                                            exact / max(cands, 1))},
        "batched": {"step_ms_sum": kernel_ms, "events": events, "candidates_per_leg": cands / max(events, 1),
                    "class_kernel_ms": [ms for ms, _ in cls_ms], "pressure_rho_0_80": pz},
    }
    if not args.no_charged:
        log("charged soft spheres ...")
        out["charged_soft_spheres"] = charged_spheres_rate(args.charged_replicas, args.charged_legs)
        log("charged soft spheres, cell veto ...")
        out["charged_soft_spheres_cell_veto"] = charged_spheres_rate(args.charged_replicas, args.charged_legs,
                                                                     far="cell_veto")
    if not args.no_charged:
        log("flexible dipoles ...")
        out["flexible_dipoles"] = flexible_dipoles_rate(64, args.charged_legs)
    if not args.no_plain:
        log("plain chain ...")
        pl = plain_chain_rate(args.plain_legs, args.plain_seconds, args.plain_equilibration)
        out["plain"] = pl
        out["batched_over_plain"] = value / pl["events_per_sec"]
    if not args.no_cpu:
        log("cpu baseline ...")
        cb = cpu_baseline(n, lengths, pos, args.cpu_seconds, args.cpu_threads)
        out["cpu_baseline"] = cb
        out["batched_over_cpu"] = value / cb["value"]
    # matched pressure: every pair of estimates at rho = 0.8 within two combined standard errors
    ests = {"batched": pz}
    if "plain" in out:
        ests["plain"] = out["plain"]["pressure"]
    if "cpu_baseline" in out:
        ests["cpu"] = out["cpu_baseline"]["pressure_rho_0_80"]
    keys, match = sorted(ests), {}
    for i, a in enumerate(keys):
        for b in keys[i + 1:]:
            ea, eb = ests[a], ests[b]
            if ea.get("mean") is None or eb.get("mean") is None:
                continue
            sig = float(np.hypot(ea["err"], eb["err"]))
            match["%s_vs_%s" % (a, b)] = dict(diff=ea["mean"] - eb["mean"], sigma=sig,
                                              within_2sigma=bool(abs(ea["mean"] - eb["mean"]) <= 2 * sig))
    out["matched_pressure"] = match
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
