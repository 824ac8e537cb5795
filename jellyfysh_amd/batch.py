"""Batched independent replicas on the GPU: host side of ecmc_batch_* in the C ABI.

Also holds the synthetic workload of the batched LJ benchmark: replica sizes drawn
log-uniformly, densities uniformly, lattice starts with a small seeded jitter.
"""
import ctypes
import json
import os

import numpy as np

from . import _lib as L

BATCH_EVENT_DTYPE = np.dtype([("kind", np.int32), ("active", np.int32), ("target", np.int32), ("factor", np.int32),
                              ("displacement", np.float64), ("x", np.float64)])

TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "gfx950.json")


class BatchDesc(ctypes.Structure):
    _fields_ = [("n_replicas", ctypes.c_int32), ("n_factors", ctypes.c_int32), ("beta", ctypes.c_double),
                ("cutoff", ctypes.c_double), ("chain_length", ctypes.c_double), ("power", ctypes.c_double * 2),
                ("prefactor", ctypes.c_double * 2), ("seed", ctypes.c_uint64)]


def lattice_positions(n, system_length, jitter, rng):
    """n distinct sites, drawn uniformly, of the smallest m^3 simple-cubic grid holding n
    units (spacing L/m), plus a uniform jitter of +-jitter per coordinate, wrapped into
    [0, L)."""
    m = int(np.ceil(n ** (1.0 / 3.0) - 1e-9))
    while m ** 3 < n:
        m += 1
    a = system_length / m
    idx = np.sort(rng.choice(m ** 3, size=n, replace=False))
    pts = np.stack([idx % m, (idx // m) % m, idx // (m * m)], axis=1).astype(np.float64) * a + 0.5 * a
    pts += rng.uniform(-jitter, jitter, pts.shape)
    pts = np.mod(pts, system_length)
    pts[pts >= system_length] = 0.0
    return pts


def batched_lj_workload(n_replicas=4096, n_min=64, n_max=2048, rho_min=0.7, rho_max=0.9, seed=2024, jitter=0.01):
    """Replica sizes log-uniform in [n_min, n_max], densities uniform in [rho_min, rho_max]."""
    rng = np.random.default_rng(seed)
    n = np.rint(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_replicas))).astype(np.int32)
    rho = rng.uniform(rho_min, rho_max, n_replicas)
    lengths = (n / rho) ** (1.0 / 3.0)
    pos = [lattice_positions(int(k), float(l), jitter, rng) for k, l in zip(n, lengths)]
    return n, lengths, pos


def default_classes():
    """Size classes and wavefronts per replica from the autotuner's table when present."""
    if os.path.exists(TUNING_FILE):
        with open(TUNING_FILE) as f:
            t = json.load(f)
        return list(t["class_max_n"]), list(t["class_wpr"])
    return [256, 768, 16384], [1, 2, 4]


class DeviceBatch:
    """Independent replicas resident on the device (ecmc_batch_*)."""

    def __init__(self, n_per_replica, system_lengths, positions, beta=1 / 1.5, cutoff=2.5, chain_length=1.0,
                 factors=((12.0, 4.0), (6.0, -4.0)), seed=1, classes=None):
        L.require_device()
        n = np.ascontiguousarray(np.asarray(n_per_replica, dtype=np.int32))
        lengths = np.ascontiguousarray(np.asarray(system_lengths, dtype=np.float64))
        if isinstance(positions, (list, tuple)):
            positions = np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, 3) for p in positions])
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
        if pos.shape[0] != int(n.sum()):
            raise ValueError("positions hold %d units, replicas need %d" % (pos.shape[0], int(n.sum())))
        d = BatchDesc()
        d.n_replicas = len(n)
        d.n_factors = len(factors)
        d.beta, d.cutoff, d.chain_length = beta, cutoff, chain_length
        for i, (p, k) in enumerate(factors):
            d.power[i], d.prefactor[i] = p, k
        d.seed = seed
        self.desc = d
        self.n = n
        self.lengths = lengths
        self._sig()
        h = ctypes.c_void_p()
        pd = ctypes.POINTER(ctypes.c_double)
        L.check(self._L.ecmc_batch_create(ctypes.byref(d), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          lengths.ctypes.data_as(pd), pos.ctypes.data_as(pd), ctypes.byref(h)))
        self._h = h
        if classes is None:
            classes = default_classes()
        self.set_classes(*classes)

    def _sig(self):
        self._L = L.lib()

    def close(self):
        if getattr(self, "_h", None):
            self._L.ecmc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_classes(self, class_max_n, class_wpr):
        mx = np.ascontiguousarray(np.asarray(class_max_n, dtype=np.int32))
        wp = np.ascontiguousarray(np.asarray(class_wpr, dtype=np.int32))
        mx[-1] = max(int(mx[-1]), int(self.n.max()))
        p = ctypes.POINTER(ctypes.c_int32)
        L.check(self._L.ecmc_batch_set_classes(self._h, len(mx), mx.ctypes.data_as(p), wp.ctypes.data_as(p)))
        self.classes = (list(mx), list(wp))

    def set_mode(self, mode):
        L.check(self._L.ecmc_batch_set_mode(self._h, int(mode)))

    def set_log(self, legs_per_replica):
        L.check(self._L.ecmc_batch_set_log(self._h, int(legs_per_replica)))
        self._log_legs = int(legs_per_replica)

    def run(self, legs_per_replica):
        L.check(self._L.ecmc_batch_run(self._h, int(legs_per_replica)))
        return self.last_ms()

    def run_graph(self, legs_per_replica, n_steps):
        L.check(self._L.ecmc_batch_run_graph(self._h, int(legs_per_replica), int(n_steps)))
        return self.last_ms()

    def set_max_blocks(self, max_blocks):
        """Limit each size class's persistent grid (0 = as many workgroups as fit)."""
        L.check(self._L.ecmc_batch_set_max_blocks(self._h, int(max_blocks)))

    def set_class_timing(self, on=True):
        L.check(self._L.ecmc_batch_set_class_timing(self._h, int(bool(on))))

    def class_ms(self):
        """Per size class: (kernel ms, launches) of the last run() call, timed with events on the
        class's own stream while class timing is on (graph replays are not class-timed)."""
        out = []
        for c in range(len(self.classes[0])):
            ms, k = ctypes.c_double(), ctypes.c_int32()
            L.check(self._L.ecmc_batch_class_ms(self._h, c, ctypes.byref(ms), ctypes.byref(k)))
            out.append((ms.value, k.value))
        return out

    def last_ms(self):
        ms = ctypes.c_double()
        L.check(self._L.ecmc_batch_last_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def states(self):
        r = len(self.n)
        active = np.zeros(r, np.int32)
        direction = np.zeros(r, np.int32)
        left = np.zeros(r)
        legs = np.zeros(r, np.uint64)
        est = np.zeros((r, 8))
        stats = np.zeros((r, 4), np.int64)
        L.check(self._L.ecmc_batch_states(self._h, active.ctypes.data, direction.ctypes.data, left.ctypes.data,
                                          legs.ctypes.data, est.ctypes.data, stats.ctypes.data))
        return dict(active=active, direction=direction, chain_left=left, legs=legs, est=est, stats=stats)

    def positions(self):
        out = np.zeros((int(self.n.sum()), 3))
        L.check(self._L.ecmc_batch_positions(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return np.split(out, np.cumsum(self.n)[:-1])

    def log(self):
        out = np.zeros(len(self.n) * self._log_legs, dtype=BATCH_EVENT_DTYPE)
        L.check(self._L.ecmc_batch_log(self._h, out.ctypes.data))
        return out.reshape(len(self.n), self._log_legs)

    def audit(self):
        v = ctypes.c_uint64()
        L.check(self._L.ecmc_batch_audit(self._h, ctypes.byref(v)))
        return v.value

    def snapshot(self):
        """Every replica's state and all positions as bytes (ecmc_batch_snapshot)."""
        n = ctypes.c_int64()
        L.check(self._L.ecmc_batch_snapshot(self._h, None, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        L.check(self._L.ecmc_batch_snapshot(self._h, buf, ctypes.byref(n)))
        return buf.raw

    def restore(self, snapshot):
        """Continue from a snapshot of a batch of the same description (ecmc_batch_restore)."""
        buf = ctypes.create_string_buffer(bytes(snapshot), len(snapshot))
        L.check(self._L.ecmc_batch_restore(self._h, buf, len(snapshot)))


def pressure(est):
    """beta P / rho = 1 + sum of lifting separations / total chain displacement, per replica."""
    est = np.atleast_2d(est)
    return 1.0 + est[:, 0] / est[:, 1]
