"""ctypes front end of the oracle library (test infrastructure only).

The oracle is the CPU restatement of the reference's ECMC inner loop (see
``ecmc_oracle.h`` for the file:line list it follows).  Only the tests, the smoke check and
the CPU-baseline leg of the benchmark may import this module; the product package never
does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libecmc_oracle.so")
_REF_DIR = os.path.join(_HERE, "_ref")

MATH_LIBM, MATH_CR = 0, 1
TAG_PAIR_CELLS, TAG_PAIR_ALL, TAG_CELL_BOUNDARY, TAG_SAMPLING, TAG_END_OF_CHAIN, TAG_END_OF_RUN, \
    TAG_START_OF_RUN = range(7)
EH_TWO_LEAF, EH_TWO_LEAF_BOUNDING, EH_CELL_BOUNDING, EH_CELL_VETO = 0, 1, 2, 3
TAG_SURPLUS, TAG_CELL_BOUNDING, TAG_CELL_VETO = 7, 8, 9
POT_HARD_SPHERE, POT_INVERSE_POWER, POT_LENNARD_JONES, POT_INVERSE_POWER_TRUNC, POT_MERGED_IMAGE_COULOMB, \
    POT_IPCB = range(6)
SCHED_LIST, SCHED_HEAP = 0, 1
TIES_SCHEDULER, TIES_FIRST_PUSH = 0, 1
MAX_TAGGERS = 16
MAX_LEAVES = 4
MAX_FACTOR_MAP = 8


class Tagger(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("event_handler", ctypes.c_int32), ("potential", ctypes.c_int32),
                ("bounding_potential", ctypes.c_int32), ("par", ctypes.c_double * 4), ("bpar", ctypes.c_double * 4),
                ("use_charge", ctypes.c_int32), ("n_create", ctypes.c_int32), ("create", ctypes.c_int32 * MAX_TAGGERS),
                ("n_trash", ctypes.c_int32), ("trash", ctypes.c_int32 * MAX_TAGGERS), ("n_handlers", ctypes.c_int32),
                ("value", ctypes.c_double), ("ivalue", ctypes.c_int32), ("ivalue2", ctypes.c_int32),
                ("fmap_local", ctypes.c_int32), ("fmap_n", ctypes.c_int32 * MAX_LEAVES),
                ("fmap_target", (ctypes.c_int32 * MAX_FACTOR_MAP) * MAX_LEAVES)]


class Config(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("n", ctypes.c_int32), ("system_length", ctypes.c_double),
                ("beta", ctypes.c_double), ("cells_per_side", ctypes.c_int32 * 3), ("neighbor_layers", ctypes.c_int32),
                ("max_occupants", ctypes.c_int32), ("scheduler", ctypes.c_int32), ("math_mode", ctypes.c_int32),
                ("rng_skip", ctypes.c_int32), ("ties", ctypes.c_int32), ("leaves_per_root", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("n_taggers", ctypes.c_int32), ("taggers", Tagger * MAX_TAGGERS)]


EVENT_DTYPE = np.dtype([("tagger", np.int32), ("flags", np.int32), ("active", np.int32), ("target", np.int32),
                        ("tq", np.float64), ("tr", np.float64), ("pos", np.float64, (3,)), ("ncand", np.int32),
                        ("chash", np.uint32)])


def build():
    """Compile the oracle (and the reference's own C files when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        d, i32, i64, u64, vp = ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_void_p
        pd = ctypes.POINTER(ctypes.c_double)
        sig = {
            "ora_mt_seed": (None, [vp, u64]), "ora_mt_u32": (ctypes.c_uint32, [vp]), "ora_random": (d, [vp]),
            "ora_expovariate": (d, [vp, d, i32]), "ora_uniform": (d, [vp, d, d]), "ora_randint": (i64, [vp, i64, i64]),
            "ora_py_mod": (d, [d, d]), "ora_py_divmod": (None, [d, d, pd, pd]),
            "ora_time_add": (None, [d, d, d, pd, pd]), "ora_time_sub": (d, [d, d, d, d]),
            "ora_pow": (d, [d, d, i32]), "ora_log": (d, [d, i32]),
            "ora_log_n": (None, [i64, pd, pd, i32]), "ora_pow_n": (None, [i64, pd, pd, pd, i32]),
            "ora_separation_norms": (None, [i32, i64, pd, d, pd, i32]),
            "ora_hs_displacement": (d, [i32, pd, pd, d]),
            "ora_hd_displacement": (d, [i32, pd, pd, d, d]),
            "ora_ip_displacement": (d, [i32, i32, d, pd, d, d, d, d, d, i32]),
            "ora_ip_derivative": (d, [i32, i32, d, pd, d, d, d, d, i32]),
            "ora_ip_potential": (d, [i32, pd, d, d, d, i32]),
            "ora_ipt_displacement": (d, [i32, i32, d, pd, d, d, d, d, d, d, i32]),
            "ora_lj_displacement": (d, [i32, i32, d, pd, d, d, d, i32]),
            "ora_lj_derivative": (d, [i32, i32, d, pd, d, d, i32]),
            "ora_dep_displacement": (d, [i32, i32, d, pd, d, i32, d, d, i32]),
            "ora_dep_derivative": (d, [i32, i32, d, pd, d, i32, d, i32]),
            "ora_ipcb_derivative": (d, [d, d, d, d]), "ora_ipcb_displacement": (d, [d, d, d, d, d, d]),
            "ora_ewald_create": (vp, [i32, i32, d, d]), "ora_ewald_destroy": (None, [vp]),
            "ora_ewald_derivative": (d, [vp, d, d, d]),
            "ora_cells_create": (vp, [i32, d, ctypes.POINTER(i32), i32]), "ora_cells_destroy": (None, [vp]),
            "ora_cells_count": (i32, [vp]), "ora_cells_bounds": (None, [vp, i32, pd, pd]),
            "ora_cells_position_to_cell": (i32, [vp, pd]), "ora_cells_nearby": (i32, [vp, i32, ctypes.POINTER(i32)]),
            "ora_cells_neighbor": (i32, [vp, i32, i32, i32]), "ora_cells_translate": (i32, [vp, i32, i32, d]),
            "ora_cell_bound_table": (None, [vp, vp, d, d, i32, d, d, d, pd]),
            "ora_engine_create": (vp, [vp, pd, pd]), "ora_engine_destroy": (None, [vp]),
            "ora_engine_run": (i64, [vp, i64, vp]), "ora_engine_done": (i32, [vp]),
            "ora_engine_error": (i32, [vp, ctypes.c_char_p, i32]), "ora_engine_positions": (None, [vp, pd]),
            "ora_engine_active": (i32, [vp, ctypes.POINTER(i32), pd, pd]),
            "ora_engine_estimators": (None, [vp, pd]),
            "ora_engine_last_candidates": (i32, [vp, ctypes.POINTER(i32), i32]),
            "ora_engine_rng_words": (i64, [vp]),
            "ora_config_hard_disks": (None, [vp, i32, d, d, i32, d, d, u64]),
            "ora_config_inverse_power_pairs": (None, [vp, i32, i32, d, d, i32, i32, pd, pd, d, d, d, u64, i32]),
            "ora_run_threads": (i64, [vp, pd, pd, i32, i64, pd]),
            "ora_walker_build": (i32, [vp, i32, ctypes.POINTER(i32), pd]), "ora_walker_free": (None, [vp]),
            "ora_walker_sample": (i32, [vp, vp]),
            "ora_heap_replay": (i32, [i32, ctypes.POINTER(i32), ctypes.POINTER(i32), pd, pd, i32, i32,
                                      ctypes.POINTER(i32)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def vec(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


class _WalkerC(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int32), ("total", ctypes.c_double), ("mean", ctypes.c_double),
                ("item0", ctypes.POINTER(ctypes.c_int32)), ("item1", ctypes.POINTER(ctypes.c_int32)),
                ("rate0", ctypes.POINTER(ctypes.c_double))]


class Walker:
    """Walker's alias table of the oracle (walker.py:53-130) over integer items."""

    def __init__(self, items, rates):
        it = np.ascontiguousarray(items, dtype=np.int32)
        r = vec(rates)
        self._w = _WalkerC()
        if lib().ora_walker_build(ctypes.byref(self._w), len(it), it.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                  _dp(r)) < 0:
            raise ValueError("Walker table failed (negative rate or leftover check)")
        m = self._w.m
        self.total, self.mean = self._w.total, self._w.mean
        self.item0 = np.ctypeslib.as_array(self._w.item0, (m,)).copy()
        self.item1 = np.ctypeslib.as_array(self._w.item1, (m,)).copy()
        self.rate0 = np.ctypeslib.as_array(self._w.rate0, (m,)).copy()

    def sample(self, mt):
        return lib().ora_walker_sample(ctypes.byref(self._w), mt._buf)

    def __del__(self):
        try:
            lib().ora_walker_free(ctypes.byref(self._w))
        except Exception:
            pass


class MT:
    """CPython-compatible MT19937 stream."""

    def __init__(self, seed):
        self._buf = ctypes.create_string_buffer(624 * 4 + 16)
        lib().ora_mt_seed(self._buf, seed)

    def u32(self):
        return lib().ora_mt_u32(self._buf)

    def random(self):
        return lib().ora_random(self._buf)

    def expovariate(self, lambd, math_mode=MATH_LIBM):
        return lib().ora_expovariate(self._buf, lambd, math_mode)

    def uniform(self, a, b):
        return lib().ora_uniform(self._buf, a, b)

    def randint(self, a, b):
        return lib().ora_randint(self._buf, a, b)


class Cells:
    def __init__(self, dim, system_length, cells_per_side, neighbor_layers=1):
        cps = (ctypes.c_int32 * 3)(*(list(cells_per_side) + [cells_per_side[0]] * (3 - len(cells_per_side))))
        self.dim = dim
        self._h = lib().ora_cells_create(dim, system_length, cps, neighbor_layers)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ora_cells_destroy(self._h)

    def count(self):
        return lib().ora_cells_count(self._h)

    def bounds(self, cell):
        mn, mx = np.zeros(3), np.zeros(3)
        lib().ora_cells_bounds(self._h, cell, _dp(mn), _dp(mx))
        return mn[:self.dim].copy(), mx[:self.dim].copy()

    def position_to_cell(self, p):
        p = vec(p)
        return lib().ora_cells_position_to_cell(self._h, _dp(p))

    def nearby(self, cell):
        out = (ctypes.c_int32 * 729)()
        n = lib().ora_cells_nearby(self._h, cell, out)
        return list(out[:n])

    def neighbor(self, cell, direction, positive):
        return lib().ora_cells_neighbor(self._h, cell, direction, int(positive))

    def cell_bound_table(self, system_length, alpha, fourier_cutoff, position_cutoff, prefactor=1.0,
                         target_charge=1.0, points_per_side=10, estimator_prefactor=1.5,
                         empirical_bound=float("inf")):
        """InnerPointEstimator bounds of the merged-image Coulomb derivative for every cell
        separation: array (cells, 3 directions, [upper, lower]), NaN for nearby separations."""
        h = lib().ora_ewald_create(int(fourier_cutoff), int(position_cutoff), alpha, system_length)
        try:
            out = np.zeros(self.count() * 6)
            lib().ora_cell_bound_table(self._h, h, prefactor, target_charge, int(points_per_side),
                                       estimator_prefactor, empirical_bound, system_length, _dp(out))
        finally:
            lib().ora_ewald_destroy(h)
        return out.reshape(-1, 3, 2)


class Engine:
    """One chain of the oracle engine."""

    def __init__(self, config, positions, charges=None):
        self.config = config
        pos = vec(positions)
        self.n, self.dim = pos.shape
        ch = vec(charges) if charges is not None else None
        self._h = lib().ora_engine_create(ctypes.byref(config), _dp(pos), _dp(ch) if ch is not None else None)
        self._check()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ora_engine_destroy(self._h)

    def _check(self):
        buf = ctypes.create_string_buffer(256)
        if lib().ora_engine_error(self._h, buf, 256):
            raise RuntimeError(buf.value.decode())

    def run(self, max_events, log=True):
        out = np.zeros(max_events, dtype=EVENT_DTYPE) if log else None
        k = lib().ora_engine_run(self._h, max_events, out.ctypes.data if log else None)
        self._check()
        return out[:k] if log else k

    @property
    def done(self):
        return bool(lib().ora_engine_done(self._h))

    def positions(self):
        out = np.zeros((self.n, self.dim))
        lib().ora_engine_positions(self._h, _dp(out))
        return out

    def active(self):
        d = ctypes.c_int32()
        q, r = ctypes.c_double(), ctypes.c_double()
        a = lib().ora_engine_active(self._h, ctypes.byref(d), ctypes.byref(q), ctypes.byref(r))
        return a, d.value, q.value, r.value

    def estimators(self):
        out = np.zeros(8)
        lib().ora_engine_estimators(self._h, _dp(out))
        return out

    def last_candidates(self):
        n = lib().ora_engine_last_candidates(self._h, None, 0)
        out = (ctypes.c_int32 * max(n, 1))()
        lib().ora_engine_last_candidates(self._h, out, n)
        return list(out[:n])

    def rng_words(self):
        return lib().ora_engine_rng_words(self._h)


def config_hard_disks(n, system_length, radius, cells_per_side, chain_time, end_time, seed):
    c = Config()
    lib().ora_config_hard_disks(ctypes.byref(c), n, system_length, radius, cells_per_side, chain_time, end_time, seed)
    return c


def config_inverse_power_pairs(dim, n, system_length, beta, cells_per_side, powers, prefactors, cutoff, chain_time,
                               end_time, seed, math_mode=MATH_LIBM):
    c = Config()
    p, k = vec(powers), vec(prefactors)
    lib().ora_config_inverse_power_pairs(ctypes.byref(c), dim, n, system_length, beta, cells_per_side, len(p),
                                         _dp(p), _dp(k), cutoff, chain_time, end_time, seed, math_mode)
    return c


def run_threads(config, positions, n_threads, events_per_replica, charges=None):
    pos = vec(positions)
    ch = vec(charges) if charges is not None else None
    secs = ctypes.c_double()
    total = lib().ora_run_threads(ctypes.byref(config), _dp(pos), _dp(ch) if ch is not None else None, n_threads,
                                  events_per_replica, ctypes.byref(secs))
    return total, secs.value


class BatchedConfig(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("n_factors", ctypes.c_int32), ("system_length", ctypes.c_double),
                ("beta", ctypes.c_double), ("cutoff", ctypes.c_double), ("chain_length", ctypes.c_double),
                ("power", ctypes.c_double * 2), ("prefactor", ctypes.c_double * 2), ("seed", ctypes.c_uint64),
                ("replica", ctypes.c_uint32), ("math_mode", ctypes.c_int32)]


class BatchedState(ctypes.Structure):
    _fields_ = [("active", ctypes.c_int32), ("direction", ctypes.c_int32), ("chain_left", ctypes.c_double),
                ("leg", ctypes.c_uint64), ("est", ctypes.c_double * 8)]


BATCHED_EVENT_DTYPE = np.dtype([("kind", np.int32), ("active", np.int32), ("target", np.int32),
                                ("factor", np.int32), ("displacement", np.float64), ("x", np.float64)])


class BatchedReplica:
    """One replica of the batched form, restated on the CPU (every candidate in FP64)."""

    def __init__(self, positions, system_length, replica, seed, beta=1 / 1.5, cutoff=2.5, chain_length=1.0,
                 factors=((12.0, 4.0), (6.0, -4.0)), math_mode=MATH_CR):
        c = BatchedConfig()
        c.math_mode = math_mode
        self.pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3)).copy()
        c.n = self.pos.shape[0]
        c.n_factors = len(factors)
        c.system_length = system_length
        c.beta = beta
        c.cutoff = min(cutoff, system_length / 2.0)
        c.chain_length = chain_length
        for i, (p, k) in enumerate(factors):
            c.power[i], c.prefactor[i] = p, k
        c.seed = seed
        c.replica = replica
        self.cfg = c
        self.state = BatchedState()
        self.state.chain_left = chain_length
        L = lib()
        L.ora_batched_run.restype = ctypes.c_int64
        L.ora_batched_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p,
                                      ctypes.c_int64, ctypes.c_void_p]

    def run(self, legs, log=True):
        out = np.zeros(legs, dtype=BATCHED_EVENT_DTYPE) if log else None
        k = lib().ora_batched_run(ctypes.byref(self.cfg), _dp(self.pos), ctypes.byref(self.state), legs,
                                  out.ctypes.data if log else None)
        return out[:k] if log else k


def philox(ctr, key):
    L = lib()
    L.ora_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
    C = (ctypes.c_uint32 * 4)(*ctr)
    K = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    L.ora_philox4x32_10(C, K, o)
    return list(o)


def ref_lib(name):
    """The reference's own C file compiled under _ref/ (heap, mic, ipcb), or None."""
    path = os.path.join(_REF_DIR, "libref_%s.so" % name)
    return ctypes.CDLL(path) if os.path.exists(path) else None
