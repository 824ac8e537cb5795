"""The INI front end: the reference-style configuration files build the same run as the
tagger lists the other tests use, and a run driven from an INI file commits the oracle's
events."""
import ctypes
import os

import numpy as np
import pytest

import systems as S

CONFIGS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jellyfysh_amd", "configs")


def _desc_bytes(d):
    return bytes(memoryview(d).cast("B"))


def test_hard_disk_ini_builds_the_standard_wiring():
    from jellyfysh_amd.chain import make_desc
    from jellyfysh_amd.config import RunConfig
    meta, pos = S.hard_disk_system()
    cfg = RunConfig.from_file(os.path.join(CONFIGS, "hard_disks.ini"))
    want = make_desc(2, 256, 1.0, 1.0, meta["cells_per_side"], S.wiring_from_meta(meta), 7, scheduler="list")
    assert _desc_bytes(cfg.desc(256, seed=7)) == _desc_bytes(want)


@pytest.mark.parametrize("name,cutoff", [("lennard_jones", 0.0), ("lennard_jones_truncated", 2.5)])
def test_lennard_jones_ini_builds_the_standard_wiring(name, cutoff):
    from jellyfysh_amd.chain import make_desc
    from jellyfysh_amd.config import RunConfig
    n = 4096
    L = (n / 0.8) ** (1 / 3)
    if cutoff:
        factors = [S.factor_tagger("lj_twelve", "inverse_power_truncated", [12.0, 4.0, 2.5]),
                   S.factor_tagger("lj_six", "inverse_power_truncated", [6.0, -4.0, 2.5])]
    else:
        factors = [S.factor_tagger("lj_twelve", "inverse_power", [12.0, 4.0]),
                   S.factor_tagger("lj_six", "inverse_power", [6.0, -4.0])]
    want = make_desc(3, n, L, 1 / 1.5, 6, S.standard_wiring(factors, 1.0, 1.0e12), 3, scheduler="heap")
    cfg = RunConfig.from_file(os.path.join(CONFIGS, name + ".ini"))
    assert _desc_bytes(cfg.desc(n, seed=3)) == _desc_bytes(want)


def test_unsupported_options_raise_configuration_error():
    from jellyfysh_amd.config import ConfigurationError, RunConfig
    text = open(os.path.join(CONFIGS, "hard_disks.ini")).read()
    with pytest.raises(ConfigurationError):
        RunConfig.from_string(text.replace("radius = ", "diameter = "))
    with pytest.raises(ConfigurationError):
        RunConfig.from_string(text.replace("hard_sphere_potential", "hard_dipole_potential"))
    with pytest.raises(ConfigurationError):
        RunConfig.from_string(text.replace("scheduler = list_scheduler", "scheduler = fifo_scheduler"))
    with pytest.raises(ConfigurationError):
        RunConfig.from_string(text.replace("create = nearby_sphere, cell_boundary\ntrash = nearby_sphere, cell_boundary\n"
                                           "internal_state_label = single_active_cell_occupancy\n"
                                           "event_handler = cell_boundary_event_handler",
                                           "create = nearby_sphere, cell_wall\ntrash = nearby_sphere\n"
                                           "event_handler = cell_boundary_event_handler"))


@pytest.mark.gpu
def test_mediator_from_ini_commits_the_oracle_events(device, oracle):
    from jellyfysh_amd.config import RunConfig, SingleProcessMediator
    meta, pos = S.hard_disk_system(n=256, seed=7)
    text = open(os.path.join(CONFIGS, "hard_disks.ini")).read().replace("end_of_run_time = 1.0e9",
                                                                         "end_of_run_time = 40.0")
    cfg = RunConfig.from_string(text)
    print("mediator from ini: hard disks, end of run at t = 40", flush=True)
    med = SingleProcessMediator(cfg, pos, seed=7)
    legs = med.run()
    meta["end_time"] = 40.0
    ref = oracle.Engine(S.oracle_config(meta, math_mode=oracle.MATH_CR, scheduler="list"), pos)
    want = ref.run(legs + 10)
    assert len(want) == legs
    assert np.array_equal(med.chain.positions(), ref.positions())


def test_device_chain_checks_charges_before_any_device_call():
    """The charges reach ecmc_system_create as n doubles: a wrong shape, or no charges for a
    wiring that multiplies by them, raises on the host before the library reads them."""
    from jellyfysh_amd.chain import DeviceChain, make_desc
    meta, taggers, pos, charges = S.coulomb_cell_bounded_system(n=8)
    desc = make_desc(3, 8, meta["system_length"], meta["beta"], meta["cells_per_side"], taggers, 1,
                     scheduler="heap", max_occupants=1)
    with pytest.raises(ValueError, match="shape"):
        DeviceChain(desc, pos, charges[:5])
    with pytest.raises(ValueError, match="no charges"):
        DeviceChain(desc, pos, None)


@pytest.mark.skipif(not os.path.isdir("/root/reference/jellyfysh/config_files"), reason="reference tree absent")
@pytest.mark.parametrize("name", ["cell_veto", "cell_bounded", "power_bounded"])
def test_coulomb_atoms_inis_keep_the_reference_tag_order(name):
    """The create and trash lists of the committed two-charge INI files, section by section, in
    the order of the reference's own files (the order of a create list sets which random words
    each candidate draws)."""
    def lists(path):
        return [ln.strip() for ln in open(path) if ln.startswith(("create", "trash"))]
    ref = "/root/reference/jellyfysh/config_files/2018_JCP_149_064113/coulomb_atoms/%s.ini" % name
    assert lists(os.path.join(CONFIGS, "coulomb_atoms_%s.ini" % name)) == lists(ref)


def test_batched_ini_is_the_benchmark_workload():
    """jellyfysh_amd/configs/batched_lennard_jones.ini describes the benchmark's batched
    workload: the same replica sizes, boxes and starts as batched_lj_workload(4096, seed=2024),
    and DeviceBatch's potential, beta, chain length and key seed 17."""
    from jellyfysh_amd.batch import batched_lj_workload
    from jellyfysh_amd.config import BatchedRunConfig, ConfigurationError
    path = os.path.join(CONFIGS, "batched_lennard_jones.ini")
    cfg = BatchedRunConfig.from_file(path)
    assert cfg.number_of_replicas == 4096
    assert cfg.batch_arguments() == dict(beta=1 / 1.5, cutoff=2.5, chain_length=1.0,
                                         factors=((12.0, 4.0), (6.0, -4.0)), seed=17)
    cfg.number_of_replicas = 64
    n, lengths, pos = cfg.workload()
    n2, l2, p2 = batched_lj_workload(64, seed=2024)
    assert np.array_equal(n, n2) and np.array_equal(lengths, l2)
    assert all(np.array_equal(a, b) for a, b in zip(pos, p2))
    text = open(path).read()
    import configparser
    cp = configparser.ConfigParser()
    cp.read_string(text.replace("power = 12", "power = 9"))
    with pytest.raises(ConfigurationError):
        BatchedRunConfig(cp)
