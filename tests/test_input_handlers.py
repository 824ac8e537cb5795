"""Input handlers (jellyfysh_amd/input_handlers.py) and the random-stream hand-over to the run.

Pins: the PDB reader reproduces the positions and charges the reference's own PDB tests expect
from the reference's test files (unittests/test_input_output_handler/test_input_handler/
test_pdb_input_handler.py, files copied under tests/golden/pdb), and raises where those tests
expect errors; the random input handler's positions are CPython ``random`` draws; after
``rng_skip`` words the oracle's stream continues exactly where CPython's does after the
placement; on the device, a chain created with the same skip commits the oracle's events."""
import ctypes
import os
import random

import numpy as np
import pytest

import systems as S

PDB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pdb")


def _pdb(name, charges=(1.0,), dim=3, L=1.0):
    from jellyfysh_amd.input_handlers import ChargeValues, PdbInputHandler
    return PdbInputHandler(os.path.join(PDB, name), [ChargeValues(list(charges), "electric_charge")], dim, L)


@pytest.mark.parametrize("name", ["atom_test_origin_0.pdb", "atom_test_origin_l2.pdb"])
def test_pdb_point_masses_as_the_reference_tests_expect(name):
    pos, charges = _pdb(name).read_point_masses()
    want = [[0.983, 0.032, 0.714], [0.618, 0.697, 0.848], [0.312, 0.806, 0.642], [0.012, 0.923, 0.100]]
    assert pos.shape == (4, 3)
    assert np.max(np.abs(pos - np.array(want))) < 5e-7            # places=6 in the reference's test
    assert np.array_equal(charges["electric_charge"], np.ones(4))
    # the coordinates pass through float32, as MDAnalysis stores them (then % L for the
    # file whose origin is at -L/2)
    first = {"atom_test_origin_0.pdb": 0.983, "atom_test_origin_l2.pdb": -0.017}[name]
    assert pos[0, 0] == float(np.float32(first)) % 1.0


def test_pdb_two_dimensions():
    pos, _ = _pdb("atom_test_2D.pdb", dim=2).read_point_masses()
    assert pos.shape == (2, 2)
    assert np.max(np.abs(pos - np.array([[0.983, 0.032], [0.618, 0.697]]))) < 5e-7


def test_pdb_composite_objects_are_read_per_residue():
    from jellyfysh_amd.input_handlers import ChargeValues, PdbInputHandler
    h = PdbInputHandler(os.path.join(PDB, "water_test_origin_0.pdb"),
                        [ChargeValues([0.41, -0.82, 0.41], "electric_charge"),
                         ChargeValues([0.0, 1.0, 0.0], "oxygen_indicator")], 3, 10.261)
    pos, charges = h.read()
    assert pos.shape == (4, 3, 3) and h.number_of_node_levels == 2
    assert np.max(np.abs(pos[0] - np.array([[7.859, 5.311, 0.879], [8.663, 5.863, 1.139],
                                            [8.794, 6.035, 2.135]]))) < 5e-6
    assert np.max(np.abs(pos[1] - np.array([[9.969, 1.519, 9.446], [10.068, 1.315, 0.160],
                                            [0.059, 1.032, 0.408]]))) < 5e-6
    assert list(charges["electric_charge"][0]) == [0.41, -0.82, 0.41]
    assert list(charges["oxygen_indicator"][1]) == [0.0, 1.0, 0.0]
    from jellyfysh_amd.output import ConfigurationError
    with pytest.raises(ConfigurationError):
        h.read_point_masses()


def test_pdb_errors_as_the_reference_raises_them():
    from jellyfysh_amd.input_handlers import ChargeValues, PdbInputHandler
    from jellyfysh_amd.output import ConfigurationError
    with pytest.raises(ConfigurationError):            # 2-D run, box with a third length
        _pdb("atom_test_2D_wrong_system_box.pdb", dim=2)
    with pytest.raises(AssertionError):                # 2-D run, position with a z component
        _pdb("atom_test_2D_wrong_positions.pdb", dim=2).read()
    with pytest.raises(OSError):
        _pdb("no_such_file.pdb")
    with pytest.raises(ConfigurationError):
        _pdb("atom_test_origin_0.pdb", dim=4)
    with pytest.raises(ConfigurationError):
        _pdb("atom_test_origin_0.pdb", L=2.0)
    with pytest.raises(ConfigurationError):
        PdbInputHandler(os.path.join(PDB, "atom_test_origin_0.pdb").replace(".pdb", ".pdbx"), [], 3, 1.0)
    with pytest.raises(ConfigurationError):
        PdbInputHandler(os.path.join(PDB, "test_non_equal_composite_point_objects.pdb"), [], 3, 1.0)
    with pytest.raises(ConfigurationError):            # a charge for each child
        _pdb("atom_test_origin_0.pdb", charges=(1.0, -1.0))
    with pytest.raises(ConfigurationError):            # charge names must differ
        PdbInputHandler(os.path.join(PDB, "atom_test_origin_0.pdb"),
                        [ChargeValues([1.0], "q"), ChargeValues([2.0], "q")], 3, 1.0)


def test_random_input_draws_cpython_random_and_hands_the_stream_on(oracle):
    from jellyfysh_amd.input_handlers import AtomRandomNodeCreator, ChargeValues, RandomInputHandler
    h = RandomInputHandler(5, AtomRandomNodeCreator([ChargeValues([1.0], "electric_charge")]))
    rng = random.Random(11)
    pos, charges, words = h.read(rng, 3, 2.5)
    ref = random.Random(11)
    assert np.array_equal(pos, np.array([[ref.uniform(0.0, 2.5) for _ in range(3)] for _ in range(5)]))
    assert words == 30 and list(charges["electric_charge"]) == [1.0] * 5
    # the oracle's stream, seeded alike and advanced by `words`, continues as CPython's does
    lib = oracle.lib()
    state = ctypes.create_string_buffer(4 * 624 + 64)
    lib.ora_mt_seed(state, 11)
    for _ in range(words):
        lib.ora_mt_u32(state)
    assert lib.ora_random(state) == rng.random()


def test_ini_random_input_handler():
    from jellyfysh_amd.config import RunConfig
    text = open(os.path.join(os.path.dirname(PDB), "..", "..", "jellyfysh_amd", "configs", "hard_disks.ini")).read()
    text += ("\n[InputOutputHandler]\ninput_handler = random_input_handler\n\n"
             "[RandomInputHandler]\nrandom_node_creator = atom_random_node_creator\nnumber_of_root_nodes = 7\n\n"
             "[AtomRandomNodeCreator]\ncharge_values = electric_charge_values (charge_values)\n\n"
             "[ElectricChargeValues]\ncharge_name = electric_charge\ncharge_values = -1\n")
    cfg = RunConfig.from_string(text)
    pos, q, words = cfg.initial_state(seed=3)
    ref = random.Random(3)
    assert np.array_equal(pos, np.array([[ref.uniform(0.0, 1.0) for _ in range(2)] for _ in range(7)]))
    assert list(q) == [-1.0] * 7 and words == 28
    assert cfg.desc(7, seed=3, rng_skip=words).rng_skip == 28


@pytest.mark.gpu
def test_gpu_chain_continues_the_stream_after_random_placement(device, oracle):
    from jellyfysh_amd.chain import DeviceChain, make_desc
    from jellyfysh_amd.input_handlers import AtomRandomNodeCreator, ChargeValues, RandomInputHandler
    n, seed = 12, 3
    meta, taggers, _, charges = S.coulomb_cell_bounded_system(n=n, seed=seed)
    h = RandomInputHandler(n, AtomRandomNodeCreator([ChargeValues([1.0], "electric_charge")]))
    pos, _, words = h.read(random.Random(seed), 3, meta["system_length"])
    desc = make_desc(3, n, meta["system_length"], meta["beta"], meta["cells_per_side"], taggers, seed,
                     scheduler="heap", max_occupants=meta["max_occupants"], rng_skip=words)
    print("random placement, then the run: cell-bounded Coulomb, n", n, flush=True)
    got = DeviceChain(desc, pos, charges).run(600)
    cfg = S.oracle_config_from_wiring(3, n, meta["system_length"], meta["beta"], meta["cells_per_side"], taggers,
                                      seed, "heap", oracle.MATH_CR, meta["max_occupants"])
    cfg.rng_skip = words
    want = oracle.Engine(cfg, pos, charges).run(600)
    assert len(got) == len(want) == 600
    for f in ("tagger", "flags", "target", "ncand"):
        assert np.array_equal(got[f], want[f]), f
    t_got, t_want = got["tq"] + got["tr"], want["tq"] + want["tr"]
    assert np.max(np.abs(t_got - t_want) / np.maximum(np.abs(t_want), 1.0)) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_dipole_creator_invariants_and_stream_position(dim):
    """DipoleRandomNodeCreator (dipole_random_node_creator.py:69-99): the atoms' minimum-image
    separation lies in [min, max]; the charges follow the atom order; the run's random stream
    continues right after the reported words, which counts the rejected direction draws."""
    from jellyfysh_amd.input_handlers import ChargeValues, DipoleRandomNodeCreator, RandomInputHandler
    L = 3.0
    h = RandomInputHandler(300, DipoleRandomNodeCreator(0.1, 0.5, [ChargeValues([0.5, -0.5], "electric_charge")]))
    rng = random.Random(21)
    pos, charges, words = h.read(rng, dim, L)
    assert pos.shape == (300, 2, dim) and pos.min() >= 0.0 and pos.max() < L
    s = pos[:, 0] - pos[:, 1]
    s -= L * np.round(s / L)
    sep = np.sqrt((s * s).sum(1))
    assert sep.min() >= 0.1 * (1 - 1e-12) and sep.max() <= 0.5 * (1 + 1e-12)
    assert charges["electric_charge"].shape == (300, 2)
    assert np.all(charges["electric_charge"] == np.array([0.5, -0.5]))
    ref = random.Random(21)
    for _ in range(words):
        ref.getrandbits(32)
    assert ref.random() == rng.random()
    # centre, direction and separation take (2 dim + 1) uniforms at least; the rejection loop
    # (acceptance pi/4 in 2D, pi/6 in 3D) adds more
    assert words > 300 * 2 * (2 * dim + 1)
