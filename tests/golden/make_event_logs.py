"""Write event-log fixtures by running the reference JeLLyFysh (pure-Python hot path) under
CPython in this container.

Why: the reference's own tests pin the potentials, cells, Time and vectors, but nothing in
them pins the sequence of events (mediator loop, TagActivator create/trash, cell occupancy
order, RNG interleaving).  These fixtures do.  The reference never travels with the
repository: only the resulting inputs/outputs (.npz of plain arrays + .json parameters) are
committed.

Usage:  python tests/golden/make_event_logs.py [/path/to/reference]

Adjustments, all outside the reference's algorithm:
  * ``CuboidCells.nearby_cells`` returns the insertion order of ``_yield_nearby_cells`` with
    duplicates dropped.  CPython iterates the reference's ``set`` of id-hashed cells in an
    address-dependent order; insertion order is what an insertion-ordered set yields.
  * The initial configuration is built directly as reference ``Node``/``Particle`` objects
    (lattice plus seeded jitter); the reference has no lattice input handler.
  * Events are recorded by wrapping ``push_event``/``get_succeeding_event`` of the scheduler
    and ``insert_into_global_state`` of the state handler.
"""
import json
import os
import random
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


class _Stop(Exception):
    pass


def _reset_modules():
    for name in list(sys.modules):
        if name == "jellyfysh" or name.startswith("jellyfysh."):
            del sys.modules[name]


def run_reference(system, n_events):
    """Build the reference object graph for `system` and record `n_events` legs."""
    _reset_modules()
    sys.dont_write_bytecode = True      # the reference tree is read-only: no __pycache__ there
    sys.path.insert(0, REF)
    try:
        import jellyfysh.setting as setting
        from jellyfysh.setting.hypercubic_setting import HypercubicSetting
        from jellyfysh.base.node import Node
        from jellyfysh.base.particle import Particle
        from jellyfysh.activator.internal_state.cell_occupancy.cells.cuboid_cells import CuboidCells
        from jellyfysh.activator.internal_state.cell_occupancy.cells.cuboid_periodic_cells import \
            CuboidPeriodicCells
        from jellyfysh.activator.internal_state.single_active_cell_occupancy import SingleActiveCellOccupancy
        from jellyfysh.activator.tag_activator import TagActivator
        from jellyfysh.activator.tagger.excluded_cells_tagger import ExcludedCellsTagger
        from jellyfysh.activator.tagger.cell_boundary_tagger import CellBoundaryTagger
        from jellyfysh.activator.tagger.active_global_state_in_state_tagger import ActiveGlobalStateInStateTagger
        from jellyfysh.activator.tagger.no_in_state_tagger import NoInStateTagger
        from jellyfysh.event_handler.two_leaf_unit_event_handler import TwoLeafUnitEventHandler
        from jellyfysh.event_handler.cell_boundary_event_handler import CellBoundaryEventHandler
        from jellyfysh.event_handler.single_independent_active_periodic_direction_end_of_chain_event_handler \
            import SingleIndependentActivePeriodicDirectionEndOfChainEventHandler
        from jellyfysh.event_handler.final_time_end_of_run_event_handler import FinalTimeEndOfRunEventHandler
        from jellyfysh.event_handler.initial_chain_start_of_run_event_handler import \
            InitialChainStartOfRunEventHandler
        from jellyfysh.event_handler.fixed_interval_sampling_event_handler import FixedIntervalSamplingEventHandler
        from jellyfysh.potential.hard_sphere_potential import HardSpherePotential
        from jellyfysh.potential.inverse_power_potential import InversePowerPotential
        from jellyfysh.potential.lennard_jones_potential import LennardJonesPotential
        from jellyfysh.scheduler.list_scheduler import ListScheduler
        from jellyfysh.state_handler.tree_state_handler import TreeStateHandler
        from jellyfysh.state_handler.physical_state.tree_physical_state import TreePhysicalState
        from jellyfysh.state_handler.lifting_state.tree_lifting_state import TreeLiftingState
        from jellyfysh.input_output_handler.input_output_handler import InputOutputHandler
        from jellyfysh.input_output_handler.input_handler.input_handler import InputHandler
        from jellyfysh.input_output_handler.output_handler.dummy_output_handler import DummyOutputHandler
        from jellyfysh.mediator.single_process_mediator import SingleProcessMediator
    finally:
        sys.path.pop(0)

    def ordered_nearby_cells(self, cell):
        cache = self.__dict__.setdefault("_golden_nearby", {})
        if cell not in cache:
            cache[cell] = list(dict.fromkeys(self._yield_nearby_cells(cell)))
        return cache[cell]
    CuboidCells.nearby_cells = ordered_nearby_cells

    dim, n, L = system["dim"], system["n"], system["system_length"]
    HypercubicSetting(beta=system["beta"], dimension=dim, system_length=L)
    setting.set_number_of_root_nodes(n)
    setting.set_number_of_nodes_per_root_node(1)
    setting.set_number_of_node_levels(1)
    positions = np.asarray(system["positions"], dtype=np.float64)

    class LatticeInput(InputHandler):
        def read(self):
            return [Node(Particle([float(v) for v in p], None)) for p in positions]

    cells = CuboidPeriodicCells(cells_per_side=[system["cells_per_side"]], neighbor_layers=1)
    occupancy = SingleActiveCellOccupancy(cells, cell_level=1, maximum_number_occupants=-1)
    pair_tags = [f["tag"] for f in system["factors"]]
    pc = pair_tags + ["cell_boundary"]
    eoc = ["end_of_chain"] + pc
    sampling = ["sampling"] if system.get("sampling_interval") else []
    taggers = []
    for f in system["factors"]:
        if f["kind"] == "hard_sphere":
            pot = HardSpherePotential(radius=f["radius"])
        elif f["kind"] == "inverse_power":
            pot = InversePowerPotential(power=f["power"], prefactor=f["prefactor"])
        elif f["kind"] == "lennard_jones":
            pot = LennardJonesPotential(prefactor=f["prefactor"], characteristic_length=f["sigma"])
        taggers.append(ExcludedCellsTagger(create=pc, trash=pc, event_handler=TwoLeafUnitEventHandler(pot),
                                           number_event_handlers=n, internal_state_label="single_active_cell_occupancy",
                                           tag=f["tag"]))
    taggers.append(CellBoundaryTagger(create=pc, trash=pc, event_handler=CellBoundaryEventHandler(),
                                      internal_state_label="single_active_cell_occupancy", tag="cell_boundary"))
    if sampling:
        taggers.append(NoInStateTagger(create=sampling, trash=sampling, tag="sampling",
                                       event_handler=FixedIntervalSamplingEventHandler(
                                           sampling_interval=system["sampling_interval"],
                                           output_handler="dummy_output_handler")))
    taggers.append(ActiveGlobalStateInStateTagger(
        create=eoc, trash=eoc, tag="end_of_chain",
        event_handler=SingleIndependentActivePeriodicDirectionEndOfChainEventHandler(chain_time=system["chain_time"])))
    taggers.append(NoInStateTagger(create=["end_of_run"], trash=eoc + sampling + ["start_of_run", "end_of_run"],
                                   tag="end_of_run",
                                   event_handler=FinalTimeEndOfRunEventHandler(end_of_run_time=system["end_time"])))
    taggers.append(NoInStateTagger(create=pc + sampling + ["end_of_chain", "end_of_run"], trash=["start_of_run"],
                                   tag="start_of_run",
                                   event_handler=InitialChainStartOfRunEventHandler(
                                       initial_direction_of_motion=0, speed=1.0, initial_active_identifier=[0])))
    tag_index = {t.tag: i for i, t in enumerate(taggers)}
    activator = TagActivator(taggers, internal_states=[occupancy])
    if system["scheduler"] == "list":
        scheduler = ListScheduler()
    else:
        # The reference's heap scheduler needs its cffi-compiled C module; the list
        # scheduler returns the same earliest event whenever times are distinct.
        scheduler = ListScheduler()
    state_handler = TreeStateHandler(TreePhysicalState(), TreeLiftingState())
    io = InputOutputHandler(LatticeInput(), [DummyOutputHandler()])
    random.seed(system["seed"])
    mediator = SingleProcessMediator(io, state_handler, scheduler, activator)

    times = {}
    record = {"tagger": [], "active": [], "target": [], "tq": [], "tr": [], "pos": [], "new_active": []}
    handler_tagger = activator._event_handler_tagger_dictionary
    orig_push, orig_succ, orig_insert = scheduler.push_event, scheduler.get_succeeding_event, \
        state_handler.insert_into_global_state
    pending = {}

    def push_event(time, eh):
        times[eh] = time
        orig_push(time, eh)

    def get_succeeding_event():
        eh = orig_succ()
        active = [k[0] for k in state_handler._lifting_state._lifting_dictionary.keys()]
        pending["eh"] = eh
        pending["active"] = active[0] if active else -1
        return eh

    depth = [0]

    def insert(out_state):
        # insert_into_global_state recurses into the children through the instance
        # attribute; record only the outermost call of each leg.
        depth[0] += 1
        try:
            orig_insert(out_state)
        finally:
            depth[0] -= 1
        if depth[0]:
            return
        eh = pending["eh"]
        t = times[eh]
        tag = handler_tagger[eh].tag
        target = -1
        if isinstance(eh, TwoLeafUnitEventHandler):
            target = eh._leaf_units[eh._active_leaf_unit_index ^ 1].identifier[0]
        elif isinstance(eh, SingleIndependentActivePeriodicDirectionEndOfChainEventHandler):
            target = mediator._out_state_arguments[eh][0][0][0]
        a = pending["active"]
        pos = [0.0, 0.0, 0.0]
        if a >= 0:
            p = state_handler._physical_state.get((a,)).value.position
            pos[:len(p)] = p
        new_active = [k[0] for k in state_handler._lifting_state._lifting_dictionary.keys()]
        record["tagger"].append(tag_index[tag])
        record["active"].append(a)
        record["target"].append(target)
        record["tq"].append(t.quotient)
        record["tr"].append(t.remainder)
        record["pos"].append(pos)
        record["new_active"].append(new_active[0] if new_active else -1)
        if len(record["tagger"]) >= n_events:
            raise _Stop()

    scheduler.push_event = push_event
    scheduler.get_succeeding_event = get_succeeding_event
    state_handler.insert_into_global_state = insert
    try:
        mediator.run()
    except _Stop:
        pass
    final = np.array([state_handler._physical_state.get((i,)).value.position for i in range(n)])
    return {k: np.asarray(v) for k, v in record.items()}, final, [t.tag for t in taggers]


def fcc(cells, a, jitter, rng):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pts = np.array([(np.array([i, j, k]) + b) * a for i in range(cells) for j in range(cells) for k in range(cells)
                    for b in base])
    return np.mod(pts + rng.uniform(-jitter, jitter, pts.shape), a * cells)


def square(m, jitter_frac, rng):
    xs = (np.arange(m) + 0.5) / m
    pts = np.array([[x, y] for y in xs for x in xs])
    return np.mod(pts + rng.uniform(-1, 1, pts.shape) * jitter_frac / m, 1.0)


def systems():
    rng = np.random.default_rng(20240611)
    out = []
    # 2-D hard disks, phi = 0.60, list scheduler (hard_disk_dipoles_cells.ini wiring)
    n = 64
    radius = (0.60 / (n * np.pi)) ** 0.5
    out.append(dict(name="hard_disks_n64", dim=2, n=n, system_length=1.0, beta=1.0, cells_per_side=9,
                    factors=[dict(tag="nearby_sphere", kind="hard_sphere", radius=radius)], chain_time=0.5,
                    end_time=1.0e9, scheduler="list", seed=5, sampling_interval=0.37,
                    positions=square(8, 0.01, rng).tolist(), events=20000))
    # 3-D LJ 12-6 as two inverse-power factors (p=12, k=4 and p=6, k=-4), rho = 0.8
    n = 256
    L = (n / 0.8) ** (1 / 3)
    out.append(dict(name="lj_two_factor_n256", dim=3, n=n, system_length=L, beta=1 / 1.5, cells_per_side=4,
                    factors=[dict(tag="lj_twelve", kind="inverse_power", power=12.0, prefactor=4.0),
                             dict(tag="lj_six", kind="inverse_power", power=6.0, prefactor=-4.0)],
                    chain_time=1.0, end_time=1.0e9, scheduler="heap", seed=11, sampling_interval=0.0,
                    positions=fcc(4, L / 4, 0.01, rng).tolist(), events=1500))
    # 3-D LJ as the single Mexican-hat factor (lennard_jones_potential.py)
    n = 108
    L = (n / 0.8) ** (1 / 3)
    out.append(dict(name="lj_mexican_hat_n108", dim=3, n=n, system_length=L, beta=1 / 1.5, cells_per_side=3,
                    factors=[dict(tag="lj", kind="lennard_jones", prefactor=1.0, sigma=1.0)],
                    chain_time=1.0, end_time=1.0e9, scheduler="heap", seed=3, sampling_interval=0.0,
                    positions=fcc(3, L / 3, 0.01, rng).tolist(), events=1500))
    return out


def main():
    for system in systems():
        record, final, tags = run_reference(system, system["events"])
        meta = {k: v for k, v in system.items() if k != "positions"}
        meta["taggers"] = tags
        with open(os.path.join(OUT, system["name"] + ".json"), "w") as f:
            json.dump(meta, f, indent=1)
        np.savez_compressed(os.path.join(OUT, system["name"] + ".npz"),
                            positions=np.asarray(system["positions"]), final_positions=final, **record)
        print(system["name"], len(record["tagger"]), "events", np.bincount(record["tagger"]))


if __name__ == "__main__":
    main()
