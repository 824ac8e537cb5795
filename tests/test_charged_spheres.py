"""Charged soft spheres (jellyfysh_amd/workloads.py) and independent systems advanced in one
launch (ecmc_systems_run, jellyfysh_amd.chain.run_together).

Pins: the overlap-rejected placement keeps every minimum-image distance at or above 0.8 and
is reproducible; the soft-sphere wiring (the reference's coulomb_atoms/cell_bounded.ini for
the Coulomb factor plus an exact inverse-power-9 factor over all pairs) commits the oracle's
event sequence on the device; systems advanced together in one launch end in exactly the
state that each reaches alone (positions, estimators, random words, active unit), so the
multi-system launch inherits the single chain's parity with the oracle."""
import numpy as np
import pytest

import systems as S


def _min_image_distance(pos, L):
    s = pos[:, None, :] - pos[None, :, :]
    s -= L * np.round(s / L)
    d2 = np.einsum("ijk,ijk->ij", s, s)
    d2[np.diag_indices(len(pos))] = np.inf
    return np.sqrt(d2.min())


def test_placement_rejects_overlaps_and_is_reproducible():
    from jellyfysh_amd.workloads import charged_soft_spheres
    a = charged_soft_spheres(n=512, system_length=8.0, seed=5)
    b = charged_soft_spheres(n=512, system_length=8.0, seed=5)
    assert np.array_equal(a["positions"], b["positions"])
    assert a["positions"].min() >= 0.0 and a["positions"].max() < 8.0
    assert _min_image_distance(a["positions"], 8.0) >= 0.8
    assert a["charges"].sum() == 0.0 and np.count_nonzero(a["charges"] > 0) == 256
    c = charged_soft_spheres(n=512, system_length=8.0, seed=6)
    assert not np.array_equal(a["positions"], c["positions"])


def _small(seed):
    from jellyfysh_amd.workloads import charged_soft_spheres
    return charged_soft_spheres(n=32, system_length=4.0, seed=seed, cells_per_side=4)


def _oracle_engine(oracle, sysd, math_mode):
    cfg = S.oracle_config_from_wiring(3, sysd["n"], sysd["system_length"], sysd["beta"], sysd["cells_per_side"],
                                      sysd["taggers"], sysd["seed"], scheduler="heap", math_mode=math_mode,
                                      max_occupants=sysd["max_occupants"])
    return oracle.Engine(cfg, sysd["positions"], sysd["charges"])


def test_oracle_runs_the_soft_sphere_wiring(oracle):
    sysd = _small(1)
    log = _oracle_engine(oracle, sysd, oracle.MATH_LIBM).run(3000)
    assert len(log) == 3000
    names = [t["tag"] for t in sysd["taggers"]]
    counts = np.bincount(log["tagger"], minlength=len(names))
    # every factor commits events: Coulomb through all three candidate sets, the repulsion
    # through its exact inversion
    for tag in ("coulomb_nearby", "soft_sphere", "cell_boundary", "end_of_chain"):
        assert counts[names.index(tag)] > 0, tag


@pytest.mark.gpu
def test_gpu_soft_spheres_match_oracle(device, oracle):
    from jellyfysh_amd.workloads import make_chain
    sysd = _small(2)
    legs = 1500
    print("charged soft spheres n=32, %d legs" % legs, flush=True)
    got = make_chain(sysd).run(legs)
    want = _oracle_engine(oracle, sysd, oracle.MATH_CR).run(legs)
    assert len(got) == len(want) == legs
    for f in ("tagger", "flags", "target", "ncand"):
        bad = np.flatnonzero(got[f] != want[f])
        assert bad.size == 0, "field %s differs first at leg %d" % (f, bad[0])
    # the cell-bound tables come from the device's erfc/exp and glibc's: times to rounding
    t_got, t_want = got["tq"] + got["tr"], want["tq"] + want["tr"]
    assert np.max(np.abs(t_got - t_want) / np.maximum(np.abs(t_want), 1.0)) < 1e-12


def _state(c):
    return c.positions(), c.estimators(), c.rng_words(), c.active()


@pytest.mark.gpu
def test_gpu_systems_run_together_end_as_alone(device):
    from jellyfysh_amd.chain import DeviceChain, make_desc, run_together
    from jellyfysh_amd.workloads import make_chain
    # different systems in one launch: cell-bounded Coulomb (two seeds) and charged soft spheres
    builders = []
    for seed in (3, 4):
        meta, taggers, pos, charges = S.coulomb_cell_bounded_system(n=12, seed=seed)
        desc = make_desc(3, 12, meta["system_length"], meta["beta"], meta["cells_per_side"], taggers, seed,
                         scheduler="heap", max_occupants=1)
        builders.append(lambda desc=desc, pos=pos, charges=charges: DeviceChain(desc, pos, charges))
    builders.append(lambda: make_chain(_small(7)))
    legs = 800
    print("three systems, %d legs, together and alone" % legs, flush=True)
    together = [b() for b in builders]
    n, ms = run_together(together, legs)
    assert list(n) == [legs] * 3 and ms > 0.0
    for b, t in zip(builders, together):
        alone = b()
        assert alone.run(legs, log=False) == legs
        for x, y in zip(_state(t), _state(alone)):
            if isinstance(x, np.ndarray):
                assert np.array_equal(x, y)
            else:
                assert x == y
        # and they continue alike
        n2, _ = run_together([t], 100)
        assert alone.run(100, log=False) == n2[0] == 100
        assert np.array_equal(t.positions(), alone.positions())
    with pytest.raises(Exception):
        run_together([together[0], together[0]], 10)


@pytest.mark.gpu
def test_gpu_soft_sphere_replicas_at_size(device):
    """Sixteen N = 512 replicas in one launch; two of them replayed alone."""
    from jellyfysh_amd.chain import run_together
    from jellyfysh_amd.workloads import charged_soft_sphere_chains, make_chain, charged_soft_spheres
    legs = 300
    print("16 charged soft-sphere replicas, N = 512, %d legs" % legs, flush=True)
    chains = charged_soft_sphere_chains(16, seed=100)
    n, _ = run_together(chains, legs)
    assert list(n) == [legs] * 16
    for i in (0, 15):
        alone = make_chain(charged_soft_spheres(seed=100 + i))
        assert alone.run(legs, log=False) == legs
        assert np.array_equal(alone.positions(), chains[i].positions())
        assert np.array_equal(alone.estimators(), chains[i].estimators())


def test_charged_soft_sphere_ini_is_the_workload_wiring():
    """jellyfysh_amd/configs/charged_soft_spheres.ini, read by the INI front end, describes the
    same run (byte for byte) as the tagger list of jellyfysh_amd.workloads."""
    import os
    from jellyfysh_amd.chain import make_desc
    from jellyfysh_amd.config import RunConfig
    from jellyfysh_amd.workloads import charged_soft_spheres
    ini = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jellyfysh_amd", "configs",
                       "charged_soft_spheres.ini")
    cfg = RunConfig.from_file(ini)
    sysd = charged_soft_spheres(seed=3)
    a = cfg.desc(512, seed=3)
    b = make_desc(3, 512, sysd["system_length"], sysd["beta"], sysd["cells_per_side"], sysd["taggers"], 3,
                  scheduler="heap", max_occupants=sysd["max_occupants"])
    assert bytes(a) == bytes(b)
