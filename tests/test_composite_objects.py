"""Composite point objects (dipoles): the factor-type-map tagger, leaf-level end-of-chain draws
and the dipole placement, against the reference's own event log of a 6-dipole run
(tests/golden/dipoles_factor_maps_r6.*, written by tests/golden/make_dipole_event_log.py)."""
import numpy as np
import pytest

import systems as S

FIELDS = ["tagger", "target", "tq", "tr", "pos"]


def _fixture():
    return S.load_fixture("dipoles_factor_maps_r6")


def test_dipole_placement_is_the_reference_input_handler():
    meta, g = _fixture()
    pos, charges, words = S.dipole_words(meta)
    assert np.array_equal(pos, g["positions"])
    assert np.array_equal(charges, g["charges"])
    assert words >= 6 * meta["roots"] * meta["dim"]       # centre, direction (with rejections), separation


def test_oracle_reproduces_the_reference_dipole_log(oracle):
    meta, g = _fixture()
    n = len(g["tagger"])
    e = oracle.Engine(S.dipole_oracle_config(meta, math_mode=oracle.MATH_LIBM), g["positions"], g["charges"])
    log = e.run(n)
    assert len(log) == n
    for f in FIELDS:
        eq = log[f] == g[f]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.flatnonzero(~eq)
        assert bad.size == 0, (f, int(bad[0]))
    assert np.array_equal(log["active"], g["active"])
    assert np.array_equal(e.positions(), g["final_positions"])
    counts = np.bincount(log["tagger"], minlength=7)
    assert counts[0] > 1000 and counts[1] > 5 and counts[2] > 100 and counts[4] > 100


@pytest.mark.gpu
def test_gpu_dipoles_match_oracle_and_reference(device, oracle):
    from jellyfysh_amd.chain import DeviceChain
    meta, g = _fixture()
    n = len(g["tagger"])
    print("dipoles with factor type maps, %d legs" % n, flush=True)
    chain = DeviceChain(S.dipole_device_desc(meta), g["positions"], g["charges"])
    log = np.concatenate([chain.run(n // 2), chain.run(n - n // 2)])
    e = oracle.Engine(S.dipole_oracle_config(meta, math_mode=oracle.MATH_CR), g["positions"], g["charges"])
    ref = e.run(n)
    for f in ("tagger", "flags", "active", "target", "tq", "tr", "pos", "ncand", "chash"):
        eq = log[f] == ref[f]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.flatnonzero(~eq)
        assert bad.size == 0, (f, int(bad[0]))
    assert chain.rng_words() == e.rng_words()
    assert np.array_equal(chain.positions(), e.positions())
    # against the reference's own log: targets and taggers equal, times within 1e-12
    assert np.array_equal(log["target"], g["target"]) and np.array_equal(log["tagger"], g["tagger"])
    t, tg = log["tq"] + log["tr"], g["tq"] + g["tr"]
    assert np.max(np.abs(t - tg) / np.maximum(np.abs(tg), 1.0)) < 1e-12


def _hdd():
    return S.load_fixture("hard_disk_dipoles_r81")


def test_pdb_reader_gives_the_fixture_start():
    """The 81-dipole .pdb start through the INI front end (composite objects from a .pdb file)
    equals the leaf positions the fixture's reference run started from."""
    import os
    from jellyfysh_amd.config import RunConfig
    meta, g = _hdd()
    cfg = RunConfig.from_file(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                           "jellyfysh_amd", "configs", "hard_disk_dipoles_cells.ini"))
    pos, q, words = cfg.initial_state(0)
    assert cfg.leaves_per_root == 2 and words == 0
    assert np.array_equal(pos, g["positions"])
    assert np.array_equal(q, g["charges"])


def test_oracle_reproduces_the_reference_hard_disk_dipole_log(oracle):
    meta, g = _hdd()
    n = len(g["tagger"])
    e = oracle.Engine(S.hard_disk_dipole_oracle_config(meta, math_mode=oracle.MATH_LIBM), g["positions"])
    log = e.run(n)
    assert len(log) == n
    for f in FIELDS:
        eq = log[f] == g[f]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.flatnonzero(~eq)
        assert bad.size == 0, (f, int(bad[0]))
    assert np.array_equal(log["active"], g["active"])
    assert np.array_equal(e.positions(), g["final_positions"])


@pytest.mark.gpu
def test_gpu_hard_disk_dipoles_match_oracle_and_reference(device, oracle):
    from jellyfysh_amd.chain import DeviceChain
    meta, g = _hdd()
    n = len(g["tagger"])
    print("81 hard-disk dipoles, %d legs" % n, flush=True)
    chain = DeviceChain(S.hard_disk_dipole_device_desc(meta), g["positions"])
    log = chain.run(n)
    e = oracle.Engine(S.hard_disk_dipole_oracle_config(meta, math_mode=oracle.MATH_CR), g["positions"])
    ref = e.run(n)
    for f in ("tagger", "flags", "active", "target", "tq", "tr", "pos", "ncand", "chash"):
        eq = log[f] == ref[f]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.flatnonzero(~eq)
        assert bad.size == 0, (f, int(bad[0]))
    assert np.array_equal(chain.positions(), e.positions())
    for f in FIELDS:
        eq = log[f] == g[f]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        assert eq.all(), f
