"""One ECMC chain on the GPU: the host side of ecmc_system_* in the C ABI.

``make_desc`` builds the run configuration from a list of tagger dictionaries that mirror
the reference's INI sections (TagActivator taggers with their create/trash tag lists and
event handlers, mediator/..., activator/tag_activator.py:166-316); ``DeviceChain`` owns
the device state and advances it.
"""
import ctypes

import numpy as np

from . import _lib as L

_KINDS = {"pair_cells": L.TAG_PAIR_CELLS, "pair_all": L.TAG_PAIR_ALL, "cell_boundary": L.TAG_CELL_BOUNDARY,
          "sampling": L.TAG_SAMPLING, "end_of_chain": L.TAG_END_OF_CHAIN, "end_of_run": L.TAG_END_OF_RUN,
          "start_of_run": L.TAG_START_OF_RUN, "surplus": L.TAG_SURPLUS, "cell_bounding": L.TAG_CELL_BOUNDING,
          "cell_veto": L.TAG_CELL_VETO, "factor_map": L.TAG_FACTOR_MAP}
_PAIR_KINDS = ("pair_cells", "pair_all", "surplus", "cell_bounding", "cell_veto", "factor_map")
# InnerPointEstimator defaults (inner_point_estimator.py:45-46): prefactor, points_per_side,
# target_charge, empirical_bound
CELL_BOUNDING_DEFAULTS = (1.5, 10.0, 1.0, float("inf"))
_POTENTIALS = {"hard_sphere": L.POT_HARD_SPHERE, "inverse_power": L.POT_INVERSE_POWER,
               "lennard_jones": L.POT_LENNARD_JONES, "inverse_power_truncated": L.POT_INVERSE_POWER_TRUNC,
               "merged_image_coulomb": L.POT_MERGED_IMAGE_COULOMB,
               "inverse_power_coulomb_bounding": L.POT_INVERSE_POWER_COULOMB_BOUND,
               "displaced_even_power": L.POT_DISPLACED_EVEN_POWER, "cell_bounding": L.POT_CELL_BOUNDING,
               "hard_dipole": L.POT_HARD_DIPOLE}


def potential_code(name):
    try:
        return _POTENTIALS[name]
    except KeyError:
        raise ValueError("unknown potential %r" % (name,))


def fill_factor_map(g, factor_map, leaves_per_root):
    """The factor type map of one tagger (``{"local": bool, "targets": [[leaf numbers] per leaf
    number of the active unit]}``, two-object numbering) into an ecmc_tagger_desc / ora_tagger."""
    targets = factor_map["targets"]
    if len(targets) > L.MAX_LEAVES or len(targets) != leaves_per_root:
        raise ValueError("a factor type map lists the entries of each of the %d leaf numbers" % leaves_per_root)
    g.fmap_local = 1 if factor_map["local"] else 0
    for k, lst in enumerate(targets):
        if len(lst) > L.MAX_FACTOR_MAP:
            raise ValueError("at most %d factor type map entries per leaf number" % L.MAX_FACTOR_MAP)
        g.fmap_n[k] = len(lst)
        for e, tl in enumerate(lst):
            g.fmap_target[k][e] = int(tl)


def kernel_variant_of(desc):
    """(features the descriptor's taggers use, features of the specialized leg-loop kernel that
    would run its system): ecmc_desc_kernel_variant, no device needed."""
    f, v = ctypes.c_int32(), ctypes.c_int32()
    L.check(L.lib().ecmc_desc_kernel_variant(ctypes.byref(desc), ctypes.byref(f), ctypes.byref(v)))
    return f.value, v.value


def make_desc(dim, n, system_length, beta, cells_per_side, taggers, seed, scheduler="heap", neighbor_layers=1,
              max_occupants=-1, rng_skip=0, leaves_per_root=1):
    """Fill an ecmc_system_desc.

    Each tagger is a dict with ``tag``, ``kind`` (pair_cells, pair_all, cell_boundary,
    sampling, end_of_chain, end_of_run, start_of_run), ``create`` and ``trash`` (lists of
    tags), and per kind: ``potential``/``params`` (pair taggers, params as in the reference
    constructors), ``bounding_potential``/``bounding_params`` for the bounding handler,
    ``charge`` (bool), ``number_event_handlers``; ``sampling_interval`` and
    ``first_event_time_zero``; ``chain_time``; ``end_of_run_time``;
    ``initial_direction_of_motion``, ``speed``, ``initial_active_identifier``.  ``rng_skip``
    words of the random stream are drawn before the run (a random input handler's draws).
    Composite point objects: ``leaves_per_root`` leaf units per object (unit root * leaves + leaf),
    pair taggers of kind ``factor_map`` with a ``factor_map`` (see fill_factor_map).
    """
    d = L.SystemDesc()
    d.rng_skip = int(rng_skip)
    d.leaves_per_root = int(leaves_per_root)
    d.dim, d.n, d.system_length, d.beta = int(dim), int(n), float(system_length), float(beta)
    cps = list(cells_per_side) if np.ndim(cells_per_side) else [int(cells_per_side)]
    for i in range(3):
        d.cells_per_side[i] = int(cps[i] if i < len(cps) else cps[0])
    d.neighbor_layers = int(neighbor_layers)
    d.max_occupants = int(max_occupants)
    d.scheduler = L.SCHED_LIST if scheduler == "list" else L.SCHED_HEAP
    d.seed = int(seed)
    if len(taggers) > L.MAX_TAGGERS:
        raise ValueError("at most %d taggers" % L.MAX_TAGGERS)
    index = {t["tag"]: i for i, t in enumerate(taggers)}
    d.n_taggers = len(taggers)
    for i, t in enumerate(taggers):
        g = d.taggers[i]
        g.kind = _KINDS[t["kind"]]
        for lst, field, count in ((t.get("create", []), g.create, "n_create"), (t.get("trash", []), g.trash, "n_trash")):
            for j, tag in enumerate(lst):
                if tag not in index:
                    raise ValueError("tag %r of tagger %r does not exist as a tagger" % (tag, t["tag"]))
                field[j] = index[tag]
            setattr(g, count, len(lst))
        g.n_handlers = int(t.get("number_event_handlers", n))
        if t["kind"] in _PAIR_KINDS:
            if t["kind"] == "cell_bounding":
                g.event_handler = L.EH_CELL_BOUNDING
            elif t["kind"] == "cell_veto":
                g.event_handler = L.EH_CELL_VETO
            else:
                g.event_handler = L.EH_TWO_LEAF_BOUNDING if t.get("bounding_potential") else L.EH_TWO_LEAF
            g.potential = potential_code(t["potential"])
            for j, v in enumerate(t.get("params", ())):
                g.par[j] = float(v)
            if t["kind"] in ("cell_bounding", "cell_veto"):
                g.bounding_potential = L.POT_CELL_BOUNDING
                bp = list(t.get("bounding_params", ()))
                bp = bp + list(CELL_BOUNDING_DEFAULTS[len(bp):])
                for j, v in enumerate(bp):
                    g.bpar[j] = float(v)
            elif t.get("bounding_potential"):
                g.bounding_potential = potential_code(t["bounding_potential"])
                for j, v in enumerate(t.get("bounding_params", ())):
                    g.bpar[j] = float(v)
            g.use_charge = 1 if t.get("charge") else 0
            if t["kind"] == "factor_map":
                fill_factor_map(g, t["factor_map"], int(leaves_per_root))
        elif t["kind"] == "sampling":
            g.value = float(t["sampling_interval"])
            g.ivalue = 1 if t.get("first_event_time_zero") else 0
        elif t["kind"] == "end_of_chain":
            g.value = float(t["chain_time"])
        elif t["kind"] == "end_of_run":
            g.value = float(t["end_of_run_time"])
        elif t["kind"] == "start_of_run":
            g.value = float(t.get("speed", 1.0))
            g.ivalue = int(t.get("initial_direction_of_motion", 0))
            g.ivalue2 = int(t.get("initial_active_identifier", 0))
    return d


class DeviceChain:
    """Device-resident state of one chain plus its event loop (ecmc_system_*)."""

    def __init__(self, desc, positions, charges=None):
        self.desc = desc
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64))
        if pos.shape != (desc.n, desc.dim):
            raise ValueError("positions must have shape (n, dim) = (%d, %d)" % (desc.n, desc.dim))
        self._charges = None if charges is None else np.ascontiguousarray(np.asarray(charges, dtype=np.float64))
        if self._charges is not None and self._charges.shape != (desc.n,):
            raise ValueError("charges must have shape (n,) = (%d,)" % desc.n)
        if self._charges is None and any(desc.taggers[t].use_charge for t in range(desc.n_taggers)):
            raise ValueError("a tagger multiplies by the charges (charge = ...), but no charges were given")
        L.require_device()
        h = ctypes.c_void_p()
        pd = ctypes.POINTER(ctypes.c_double)
        L.check(L.lib().ecmc_system_create(ctypes.byref(desc), pos.ctypes.data_as(pd),
                                           None if self._charges is None else self._charges.ctypes.data_as(pd),
                                           ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().ecmc_system_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_stop_mask(self, mask):
        L.check(L.lib().ecmc_system_set_stop_mask(self._h, int(mask)))

    def run(self, max_events, log=True):
        """Commit up to max_events legs; returns the event records (or their count)."""
        n = ctypes.c_int64()
        out = np.zeros(max_events, dtype=L.EVENT_DTYPE) if log else None
        L.check(L.lib().ecmc_system_run(self._h, int(max_events), out.ctypes.data if log else None, ctypes.byref(n)))
        return out[:n.value] if log else n.value

    def positions(self):
        out = np.zeros((self.desc.n, self.desc.dim))
        L.check(L.lib().ecmc_system_positions(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def active(self):
        a, d, done = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        q, r = ctypes.c_double(), ctypes.c_double()
        L.check(L.lib().ecmc_system_active(self._h, ctypes.byref(a), ctypes.byref(d), ctypes.byref(q),
                                           ctypes.byref(r), ctypes.byref(done)))
        return a.value, d.value, q.value, r.value, bool(done.value)

    def estimators(self):
        out = np.zeros(8)
        L.check(L.lib().ecmc_system_estimators(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def rng_words(self):
        w = ctypes.c_int64()
        L.check(L.lib().ecmc_system_rng_words(self._h, ctypes.byref(w)))
        return w.value

    def launch_lds(self):
        """(dynamic LDS bytes of a launch, positions resident in LDS, ring resident in LDS)."""
        b, p, r = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        L.check(L.lib().ecmc_system_launch_lds(self._h, ctypes.byref(b), ctypes.byref(p), ctypes.byref(r)))
        return b.value, bool(p.value), bool(r.value)

    def kernel_variant(self):
        """(features the system's taggers use, features of the specialized leg-loop kernel that
        runs it): the kernel is ecmc::chain_kernel<dim, variant> alone and
        ecmc::chain_many_kernel<dim, variant> in a launch of several systems of the same variant."""
        f, v = ctypes.c_int32(), ctypes.c_int32()
        L.check(L.lib().ecmc_system_kernel_variant(self._h, ctypes.byref(f), ctypes.byref(v)))
        return f.value, v.value

    def last_kernel_ms(self):
        ms = ctypes.c_double()
        L.check(L.lib().ecmc_system_last_kernel_ms(self._h, ctypes.byref(ms)))
        return ms.value


def run_together(chains, max_events, noalloc=False):
    """Advance independent chains in one launch (ecmc_systems_run): one workgroup per chain,
    each committing up to max_events legs exactly as DeviceChain.run would (no event logs).
    With noalloc, ecmc_systems_run_noalloc launches them without allocating device memory (at
    most SYSTEMS_NOALLOC_MAX chains).  Returns the committed legs per chain and the launch's
    kernel time in milliseconds."""
    if not chains:
        raise ValueError("no chains")
    hs = (ctypes.c_void_p * len(chains))(*[c._h.value for c in chains])
    n = np.zeros(len(chains), dtype=np.int64)
    fn = L.lib().ecmc_systems_run_noalloc if noalloc else L.lib().ecmc_systems_run
    L.check(fn(hs, len(chains), int(max_events), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
    return n, chains[0].last_kernel_ms()
