"""The cell veto (cell_veto_event_handler.py:93-238, leaf_unit_cell_veto_event_handler.py:31-149,
walker.py:53-130, cell_veto_tagger.py:31-155) in the oracle and on the device.

Pins: Walker's tables of the oracle equal the two tables that the reference's own cell-veto
unit test lists for its mocked bounds (unittests/test_event_handler/
test_leaf_unit_cell_veto_event_handler.py:179-195) and a table derived by hand from walker.py's
stack steps; every item's probability mass over the table equals its rate; sampling reproduces
the rates' proportions.  The device chain commits the oracle's event sequence (sampled cells,
occupants or none, thinning) on the reference's 3 x 5 x 7 cell geometry."""
import numpy as np
import pytest

import systems as S


def _table(w):
    return [(int(a), float(r), int(b)) for a, r, b in zip(w.item0, w.rate0, w.item1)]


def _masses(w, m):
    mass = np.zeros(m)
    for a, r, b in zip(w.item0, w.rate0, w.item1):
        mass[a] += r
        if b >= 0:
            mass[b] += w.mean - r
    return mass


@pytest.mark.parametrize("rates, listing", [
    # direction 0: rates max(upper, -lower) of the mocked bounds, listed columns (small | alias)
    ([0.3, 0.7, 0.8, 1.1, 0.1, 0.2, 0.3],
     [(14, 0.3, 9), (11, 0.2, 9), (10, 0.1, 9), (9, 0.2, 8), (8, 0.5, None), (2, 0.3, 6), (6, 0.5, -1)]),
    # direction 1
    ([0.4, 0.7, 0.9, 1.0, 0.3, 0.4, 0.5],
     [(14, 0.5, 9), (11, 0.4, 9), (10, 0.3, 9), (9, 0.4, 8), (2, 0.4, 8), (8, 0.5, 6), (6, 0.6, -1)]),
])
def test_walker_tables_match_the_reference_test_listing(oracle, rates, listing):
    items = [2, 6, 8, 9, 10, 11, 14]
    w = oracle.Walker(items, rates)
    assert w.total == pytest.approx(sum(rates), rel=1e-15) and w.mean == pytest.approx(sum(rates) / 7, rel=1e-15)
    got = _table(w)
    assert len(got) == 7
    for (a, r, b), (ea, er, eb) in zip(got, listing):
        assert a == ea and r == pytest.approx(er, abs=1e-12)
        if eb is None:
            # the listing leaves this alias blank: its rate is mean - 0.5, zero up to rounding
            assert w.mean - r == pytest.approx(0.0, abs=1e-12)
        else:
            assert b == eb


def test_walker_hand_derived_table_and_masses(oracle):
    # mean 7/6; small [0, 2, 5], large [1, 3, 4]; popping from the ends gives
    # (5|4) (2|4) (4|3) (3|1) (0|1) and the leftover 1
    w = oracle.Walker(range(6), [0.5, 2.0, 0.0, 1.25, 3.0, 0.25])
    assert [(a, b) for a, _, b in _table(w)] == [(5, 4), (2, 4), (4, 3), (3, 1), (0, 1), (1, -1)]
    assert np.allclose(_masses(w, 6), [0.5, 2.0, 0.0, 1.25, 3.0, 0.25], rtol=0, atol=1e-15)
    rng = np.random.default_rng(0)
    rates = rng.exponential(1.0, 485)
    rates[rng.choice(485, 60, replace=False)] = 0.0
    w = oracle.Walker(range(485), rates)
    assert len(w.item0) == 485 and sorted(w.item0) == list(range(485))
    assert np.max(np.abs(_masses(w, 485) - rates)) < 1e-12 * rates.sum()
    with pytest.raises(ValueError):
        oracle.Walker(range(3), [0.0, 0.0, 0.0])          # leftovers fail the 1e-6 check (0/0)
    with pytest.raises(ValueError):
        oracle.Walker(range(3), [1.0, -0.5, 1.0])


def test_walker_sampling_frequencies(oracle):
    rates = [0.3, 0.7, 0.8, 1.1, 0.1, 0.2, 0.3]
    w = oracle.Walker(range(7), rates)
    mt = oracle.MT(12345)
    n = 140000
    counts = np.bincount([w.sample(mt) for _ in range(n)], minlength=7)
    p = np.array(rates) / sum(rates)
    chi2 = np.sum((counts - n * p) ** 2 / (n * p))
    assert chi2 < 30.0          # 6 degrees of freedom: p ~ 4e-5


def _oracle_engine(oracle, meta, taggers, pos, charges, math_mode):
    cfg = S.oracle_config_from_wiring(3, len(pos), meta["system_length"], meta["beta"], meta["cells_per_side"],
                                      taggers, meta["seed"], scheduler="heap", math_mode=math_mode, max_occupants=1)
    return oracle.Engine(cfg, pos, charges)


def test_oracle_cell_veto_run(oracle):
    meta, taggers, pos, charges = S.coulomb_cell_veto_system(n=12)
    e = _oracle_engine(oracle, meta, taggers, pos, charges, oracle.MATH_LIBM)
    log = e.run(20000)
    assert len(log) == 20000
    veto = log[log["tagger"] == 0]
    assert len(veto) > 0
    # sampled cells that were empty (no target) and occupied ones, lifts among the latter
    assert np.any(veto["target"] < 0) and np.any(veto["target"] >= 0)
    assert np.all(veto["target"][(veto["flags"] & 1) == 1] >= 0) and np.any(veto["flags"] & 1)
    counts = np.bincount(log["tagger"], minlength=len(taggers))
    for i, t in enumerate(taggers[:4]):
        assert counts[i] > 0, t["tag"]


@pytest.mark.gpu
@pytest.mark.parametrize("n", [2, 12])
def test_gpu_cell_veto_matches_oracle(device, oracle, n):
    from jellyfysh_amd.chain import DeviceChain, make_desc
    meta, taggers, pos, charges = S.coulomb_cell_veto_system(n=n)
    legs = 6000
    print("cell-veto chain, n = %d, %d legs" % (n, legs), flush=True)
    desc = make_desc(3, n, meta["system_length"], meta["beta"], meta["cells_per_side"], taggers, meta["seed"],
                     scheduler="heap", max_occupants=1)
    chain = DeviceChain(desc, pos, charges)
    got = chain.run(legs)
    eng = _oracle_engine(oracle, meta, taggers, pos, charges, oracle.MATH_CR)
    want = eng.run(legs)
    assert len(got) == len(want) == legs
    assert np.count_nonzero(want["tagger"] == 0) > 0
    # the event sequence is exact; the bound tables (and with them the Walker tables and the
    # candidate times) come from erfc/exp of the device's math library and of glibc
    for f in ("tagger", "flags", "target", "ncand", "chash"):
        bad = np.flatnonzero(got[f] != want[f])
        assert bad.size == 0, "field %s differs first at leg %d" % (f, bad[0])
    t_got, t_want = got["tq"] + got["tr"], want["tq"] + want["tr"]
    rel_t = np.max(np.abs(t_got - t_want) / np.maximum(np.abs(t_want), 1.0))
    rel_p = np.max(np.abs(got["pos"] - want["pos"])) / meta["system_length"]
    print("largest relative deviation: time %.2e, position %.2e" % (rel_t, rel_p), flush=True)
    assert rel_t < 1e-12 and rel_p < 1e-12
    assert chain.rng_words() == eng.rng_words()


def test_cell_veto_and_cell_bounding_sample_the_same_pressure(oracle):
    """Charged soft spheres (N = 32, L = 4, 4^3 cells) with the far cells through the cell veto
    and through the cell-bounding potential: two exact factorizations of one Boltzmann
    distribution, so the lifting estimator 1 + sum(separations) / displacement agrees.  Ten
    chunks of 1e5 legs after two of burn-in for each; the seeds are fixed, so the test is
    deterministic (1.7 combined standard errors apart with these seeds)."""
    from jellyfysh_amd.workloads import charged_soft_spheres
    means, errs = [], []
    for far in ("cell_veto", "cell_bounding"):
        s = charged_soft_spheres(n=32, system_length=4.0, seed=2, cells_per_side=4, far=far)
        cfg = S.oracle_config_from_wiring(3, 32, 4.0, 1.0, 4, s["taggers"], 2, scheduler="heap",
                                          math_mode=oracle.MATH_LIBM, max_occupants=1)
        e = oracle.Engine(cfg, s["positions"], s["charges"])
        zs, prev = [], e.estimators().copy()
        for _ in range(12):
            assert e.run(100000, log=False) == 100000
            est = e.estimators().copy()
            zs.append((est[0] - prev[0]) / (est[1] - prev[1]))
            prev = est
        zs = np.array(zs[2:])
        means.append(zs.mean())
        errs.append(zs.std(ddof=1) / np.sqrt(len(zs)))
    assert abs(means[0] - means[1]) < 3.0 * np.hypot(*errs), (means, errs)


@pytest.mark.gpu
def test_gpu_soft_spheres_with_cell_veto_match_oracle(device, oracle):
    from jellyfysh_amd.workloads import charged_soft_spheres, make_chain
    s = charged_soft_spheres(n=32, system_length=4.0, seed=5, cells_per_side=4, far="cell_veto")
    legs = 3000
    print("charged soft spheres with the cell veto, n = 32, %d legs" % legs, flush=True)
    got = make_chain(s).run(legs)
    cfg = S.oracle_config_from_wiring(3, 32, 4.0, 1.0, 4, s["taggers"], 5, scheduler="heap",
                                      math_mode=oracle.MATH_CR, max_occupants=1)
    want = oracle.Engine(cfg, s["positions"], s["charges"]).run(legs)
    assert len(got) == len(want) == legs
    assert np.count_nonzero(want["tagger"] == 0) > 0
    for f in ("tagger", "flags", "target", "ncand", "chash"):
        bad = np.flatnonzero(got[f] != want[f])
        assert bad.size == 0, "field %s differs first at leg %d" % (f, bad[0])
    t_got, t_want = got["tq"] + got["tr"], want["tq"] + want["tr"]
    assert np.max(np.abs(t_got - t_want) / np.maximum(np.abs(t_want), 1.0)) < 1e-12


def test_cell_veto_ini_is_the_workload_wiring():
    """jellyfysh_amd/configs/charged_soft_spheres_cell_veto.ini, read by the INI front end
    (cell_veto_tagger, leaf_unit_cell_veto_event_handler, inner_point_estimator), describes
    the same run, byte for byte, as jellyfysh_amd.workloads with far="cell_veto"."""
    import os
    from jellyfysh_amd.chain import make_desc
    from jellyfysh_amd.config import RunConfig
    from jellyfysh_amd.workloads import charged_soft_spheres
    ini = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jellyfysh_amd", "configs",
                       "charged_soft_spheres_cell_veto.ini")
    cfg = RunConfig.from_file(ini)
    s = charged_soft_spheres(seed=3, far="cell_veto")
    a = cfg.desc(512, seed=3)
    b = make_desc(3, 512, s["system_length"], s["beta"], s["cells_per_side"], s["taggers"], 3,
                  scheduler="heap", max_occupants=s["max_occupants"])
    assert bytes(a) == bytes(b)
    assert cfg.taggers[0]["kind"] == "cell_veto" and cfg.taggers[0]["bounding_params"] == [1.5, 10.0, 1.0,
                                                                                           float("inf")]


def _coulomb_atoms_ini(far="cell_veto"):
    """The reference's three coulomb_atoms runs as committed INI files: the cell veto
    (coulomb_atoms_cell_veto.ini), the cell-bounding potential (coulomb_atoms_cell_bounded.ini)
    and the pair through the inverse-power Coulomb bound only (coulomb_atoms_power_bounded.ini)."""
    import os
    from jellyfysh_amd.config import RunConfig
    configs = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jellyfysh_amd", "configs")
    name = {"cell_veto": "cell_veto", "cell_bounding": "cell_bounded", "power_bounded": "power_bounded"}[far]
    return RunConfig.from_file(os.path.join(configs, "coulomb_atoms_%s.ini" % name))


@pytest.mark.gpu
def test_gpu_coulomb_atoms_ini_matches_oracle(device, oracle):
    """The two-charge cell-veto INI with its random start: the device chain against the oracle,
    event for event, after the input handler's words (rng_skip)."""
    from jellyfysh_amd.chain import DeviceChain
    cfg = _coulomb_atoms_ini()
    pos, q, skip = cfg.initial_state(11)
    legs = 4000
    print("coulomb atoms with the cell veto from the INI, %d legs" % legs, flush=True)
    got = DeviceChain(cfg.desc(2, 11, rng_skip=skip), pos, q).run(legs)
    oc = S.oracle_config_from_wiring(3, 2, cfg.system_length, cfg.beta, cfg.cells_per_side, cfg.taggers, 11,
                                     cfg.scheduler, oracle.MATH_CR, cfg.max_occupants)
    oc.rng_skip = skip
    want = oracle.Engine(oc, pos, q).run(legs)
    assert len(got) == len(want) == legs
    for f in ("tagger", "flags", "target", "ncand"):
        bad = np.flatnonzero(got[f] != want[f])
        assert bad.size == 0, "field %s differs first at leg %d" % (f, bad[0])


@pytest.mark.gpu
@pytest.mark.parametrize("far", ["cell_veto", "cell_bounding", "power_bounded"])
def test_gpu_coulomb_atoms_sample_the_reference_distribution(device, far):
    """Two +1 charges (coulomb_atoms_cell_veto.ini) in 256 replicas (seeds 0-255) advanced
    together, each stopping at its own sampling events (every 0.56789): the empirical
    cumulative distribution of the sampled separations, after 20 samples of burn-in per
    replica, against the reference's reversible-Monte-Carlo data
    (tests/golden/coulomb_atoms_reference_cdf.dat), as the reference's plotting script
    compares its cell-veto run with it.  ~97 000 samples: the Kolmogorov distance of
    independent samples would be ~0.005 (measured: 0.0032); the bound 0.012 leaves room for
    their correlation."""
    import os
    from jellyfysh_amd.chain import DeviceChain, run_together
    cfg = _coulomb_atoms_ini(far)
    assert cfg.taggers[0]["kind"] == {"power_bounded": "pair_all"}.get(far, far)
    samp = [i for i, t in enumerate(cfg.taggers) if t["kind"] == "sampling"][0]
    chains = []
    for seed in range(256):
        pos, q, skip = cfg.initial_state(seed)
        c = DeviceChain(cfg.desc(2, seed, rng_skip=skip), pos, q)
        c.set_stop_mask(1 << samp)
        chains.append(c)
    L = cfg.system_length
    seps = []
    print("256 replicas of two charges (far cells: %s), 400 sampling stops" % far, flush=True)
    for it in range(400):
        n, _ = run_together(chains, 200000)
        assert np.all((n > 0) & (n < 200000))          # every replica stopped at a sampling event
        if it < 20:
            continue
        for c in chains:
            p = c.positions()
            s = np.mod((p[1] - p[0]) + L / 2, L) - L / 2
            seps.append(np.sqrt(np.sum(s * s)))
        if it % 100 == 0:
            print("stop %d" % it, flush=True)
    for c in chains:
        c.close()
    ref = np.loadtxt(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                  "coulomb_atoms_reference_cdf.dat"))
    seps = np.sort(np.array(seps))
    F = np.searchsorted(seps, ref[:, 0], side="right") / len(seps)
    ks = np.max(np.abs(F - ref[:, 1]))
    print("samples %d, Kolmogorov distance to the reference data %.4f, median %.4f" %
          (len(seps), ks, np.median(seps)), flush=True)
    assert ks < 0.012


@pytest.mark.parametrize("cps, L", [((3, 5, 7), 1.0), ((8, 8, 8), 8.0), ((6, 6, 6), 17.2355)])
def test_translate_is_the_identifier_sum(oracle, cps, L):
    """The device places the sampled cell by adding per-axis cell identifiers modulo the cells
    per side; the oracle restates the reference's translate (the cell centre shifted by the
    relative cell's lower corner, corrected into the box, then position_to_cell).  Both agree
    for every (cell, relative cell) pair."""
    import ctypes
    lib = oracle.lib()
    c = lib.ora_cells_create(3, L, (ctypes.c_int32 * 3)(*cps), 1)
    try:
        n = lib.ora_cells_count(c)
        cum = (1, cps[0], cps[0] * cps[1])
        for cell in range(n):
            a = [(cell // cum[d]) % cps[d] for d in range(3)]
            for rel in range(n):
                r = [(rel // cum[d]) % cps[d] for d in range(3)]
                want = sum(((a[d] + r[d]) % cps[d]) * cum[d] for d in range(3))
                assert lib.ora_cells_translate(c, cell, rel, L) == want, (cell, rel)
    finally:
        lib.ora_cells_destroy(c)
