"""Shared builders for test systems: fixture loading and oracle configurations."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_fixture(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        meta = json.load(f)
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    return meta, data


def oracle_config(meta, math_mode=0, scheduler=None):
    """Oracle configuration mirroring the tagger list the fixture script wired into the
    reference: factors..., cell_boundary, [sampling], end_of_chain, end_of_run, start_of_run."""
    from oracle import oracle as O
    c = O.Config()
    c.dim, c.n, c.system_length, c.beta = meta["dim"], meta["n"], meta["system_length"], meta["beta"]
    for d in range(3):
        c.cells_per_side[d] = meta["cells_per_side"]
    c.neighbor_layers = 1
    c.max_occupants = -1
    sched = scheduler or meta["scheduler"]
    c.scheduler = O.SCHED_LIST if sched == "list" else O.SCHED_HEAP
    c.math_mode = math_mode
    c.seed = meta["seed"]
    tags = meta["taggers"]
    idx = {t: i for i, t in enumerate(tags)}
    nf = len(meta["factors"])
    pc = [idx[f["tag"]] for f in meta["factors"]] + [idx["cell_boundary"]]
    eoc = [idx["end_of_chain"]] + pc
    samp = [idx["sampling"]] if "sampling" in idx else []
    c.n_taggers = len(tags)

    def lists(t, create, trash):
        t.n_create = len(create)
        for i, v in enumerate(create):
            t.create[i] = v
        t.n_trash = len(trash)
        for i, v in enumerate(trash):
            t.trash[i] = v

    for k, f in enumerate(meta["factors"]):
        t = c.taggers[idx[f["tag"]]]
        t.kind, t.event_handler, t.n_handlers = O.TAG_PAIR_CELLS, O.EH_TWO_LEAF, meta["n"]
        if f["kind"] == "hard_sphere":
            t.potential = O.POT_HARD_SPHERE
            t.par[0] = f["radius"]
        elif f["kind"] == "inverse_power":
            t.potential = O.POT_INVERSE_POWER
            t.par[0], t.par[1] = f["power"], f["prefactor"]
        elif f["kind"] == "inverse_power_truncated":
            t.potential = O.POT_INVERSE_POWER_TRUNC
            t.par[0], t.par[1], t.par[2] = f["power"], f["prefactor"], f["cutoff"]
        elif f["kind"] == "lennard_jones":
            t.potential = O.POT_LENNARD_JONES
            t.par[0], t.par[1] = f["prefactor"], f["sigma"]
        lists(t, pc, pc)
    t = c.taggers[idx["cell_boundary"]]
    t.kind, t.n_handlers = O.TAG_CELL_BOUNDARY, 1
    lists(t, pc, pc)
    if samp:
        t = c.taggers[idx["sampling"]]
        t.kind, t.n_handlers, t.value = O.TAG_SAMPLING, 1, meta["sampling_interval"]
        lists(t, samp, samp)
    t = c.taggers[idx["end_of_chain"]]
    t.kind, t.n_handlers, t.value = O.TAG_END_OF_CHAIN, 1, meta["chain_time"]
    lists(t, eoc, eoc)
    t = c.taggers[idx["end_of_run"]]
    t.kind, t.n_handlers, t.value = O.TAG_END_OF_RUN, 1, meta["end_time"]
    lists(t, [idx["end_of_run"]], eoc + samp + [idx["start_of_run"], idx["end_of_run"]])
    t = c.taggers[idx["start_of_run"]]
    t.kind, t.n_handlers, t.value, t.ivalue, t.ivalue2 = O.TAG_START_OF_RUN, 1, 1.0, 0, 0
    lists(t, pc + samp + [idx["end_of_chain"], idx["end_of_run"]], [idx["start_of_run"]])
    return c


def fcc_lattice(cells, a, jitter, rng):
    """FCC lattice of cells^3 conventional cells with lattice constant a, plus a uniform
    jitter of +-jitter per coordinate, wrapped into the box."""
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pts = np.array([(np.array([i, j, k]) + b) * a for i in range(cells) for j in range(cells) for k in range(cells)
                    for b in base])
    return np.mod(pts + rng.uniform(-jitter, jitter, pts.shape), a * cells)


def square_lattice(m, system_length, jitter_frac, rng):
    xs = (np.arange(m) + 0.5) / m * system_length
    pts = np.array([[x, y] for y in xs for x in xs])
    return np.mod(pts + rng.uniform(-1, 1, pts.shape) * jitter_frac * system_length / m, system_length)
