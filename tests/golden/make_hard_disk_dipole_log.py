"""Write the event-log fixture of the reference's hard-disk-dipole run
(config_files/hard_disk_dipoles/hard_disk_dipoles_cells.ini) by running the reference JeLLyFysh
(pure-Python hot path) under CPython in this container.

Why: it pins cells holding the leaf units of composite objects (SingleActiveCellOccupancy with
cell_level 2, ExcludedCellsTagger between leaf units), the hard-dipole tether through the factor
type map (potential/hard_dipole_potential.py) and the leaf-level end-of-chain draws, none of which
the reference's own tests follow through a run.  The reference never travels with the
repository: only the resulting inputs/outputs (.npz of plain arrays + .json parameters) are
committed.

The run: the 81 dipoles of the reference's .pdb start (read by this repository's PDB reader;
the reference reads it through MDAnalysis, which is absent here, so the script builds the same
tree: leaf positions corrected into the box, the composite object at the barycentre of its leaves
taken at the images closest to the first, as pdb_input_handler.py:137-191 does), 13 x 13 cells,
hard disks of radius 0.476190 between the leaf units of the nearby cells, the tether [0.952381,
1.047619] within a dipole, chains of 1.0, list scheduler, sampling every 10.01 from time 0.

Usage:  python tests/golden/make_hard_disk_dipole_log.py [/path/to/reference]

Adjustments, all outside the reference's algorithm:
  * ``CuboidCells.nearby_cells`` returns the insertion order of ``_yield_nearby_cells`` with
    duplicates dropped (the order of an insertion-ordered set), and the factor type map tagger's
    set of factors is taken in insertion order (as PyPy iterates sets).
  * Events are recorded by wrapping ``push_event``/``get_succeeding_event`` of the scheduler
    and ``insert_into_global_state`` of the state handler.
"""
import json
import os
import random
import sys
import tempfile

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(OUT))
PDB = os.path.join(ROOT, "jellyfysh_amd", "configs", "hard_disk_dipoles",
                   "81dipoles_min0.952380952380952_max1.047619047619048.pdb")

FACTOR_FILE = """[0, 1], Dipole
[0, 2], Sphere
[0, 3], Sphere
[1, 2], Sphere
[1, 3], Sphere
"""

SYSTEM = dict(name="hard_disk_dipoles_r81", dim=2, roots=81, leaves=2, system_length=12.836, beta=1.0, seed=29,
              cells_per_side=13, radius=0.476190476190476, minimum_separation=0.952380952380952,
              maximum_separation=1.047619047619048, chain_time=1.0, sampling_interval=10.01, end_time=1.0e9,
              scheduler="list", events=12000, charges=[1.0, -1.0])


class _Stop(Exception):
    pass


def run_reference(system):
    for name in list(sys.modules):
        if name == "jellyfysh" or name.startswith("jellyfysh."):
            del sys.modules[name]
    sys.dont_write_bytecode = True
    sys.path.insert(0, ROOT)
    from jellyfysh_amd.input_handlers import ChargeValues as OwnChargeValues, PdbInputHandler as OwnPdb
    sys.path.pop(0)
    own = OwnPdb(PDB, [OwnChargeValues(system["charges"], "electric")], system["dim"], system["system_length"])
    leaf_pos, _ = own.read()
    sys.path.insert(0, REF)
    try:
        import jellyfysh.setting as setting
        from jellyfysh.setting.hypercubic_setting import HypercubicSetting
        from jellyfysh.base.node import Node, yield_leaf_nodes
        from jellyfysh.base.particle import Particle
        from jellyfysh.activator.internal_state.cell_occupancy.cells.cuboid_cells import CuboidCells
        from jellyfysh.activator.internal_state.cell_occupancy.cells.cuboid_periodic_cells import \
            CuboidPeriodicCells
        from jellyfysh.activator.internal_state.single_active_cell_occupancy import SingleActiveCellOccupancy
        from jellyfysh.activator.tag_activator import TagActivator
        from jellyfysh.activator.tagger.excluded_cells_tagger import ExcludedCellsTagger
        from jellyfysh.activator.tagger.cell_boundary_tagger import CellBoundaryTagger
        from jellyfysh.activator.tagger.factor_type_maps import FactorTypeMaps
        from jellyfysh.activator.tagger.factor_type_map_in_state_tagger import FactorTypeMapInStateTagger
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
        from jellyfysh.potential.hard_dipole_potential import HardDipolePotential
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

    def ordered_identifiers(self, extracted_active_global_state):
        yield from dict.fromkeys(factor for root_cnode in extracted_active_global_state
                                 for leaf_cnode in yield_leaf_nodes(root_cnode)
                                 for factor in self._factor_type_map.yield_factor_identifier(
                                     leaf_cnode.value.identifier))
    FactorTypeMapInStateTagger.yield_identifiers_send_event_time = ordered_identifiers

    P, R = system["leaves"], system["roots"]
    HypercubicSetting(beta=system["beta"], dimension=system["dim"], system_length=system["system_length"])
    setting.set_number_of_root_nodes(R)
    setting.set_number_of_nodes_per_root_node(P)
    setting.set_number_of_node_levels(2)

    class PdbTree(InputHandler):
        """pdb_input_handler.py:137-191 on positions read without MDAnalysis."""
        def read(self):
            nodes = [Node() for _ in range(R)]
            for r in range(R):
                for j in range(P):
                    position = [float(v) for v in leaf_pos[r, j]]
                    setting.periodic_boundaries.correct_position(position)
                    nodes[r].add_child(Node(Particle(position=position,
                                                     charge={"electric": system["charges"][j]})))
            for node in nodes:
                first_position = node.children[0].value.position
                center_position = [entry * node.children[0].weight for entry in first_position]
                for k in range(1, len(node.children)):
                    other_position = node.children[k].value.position
                    shortest = setting.periodic_boundaries.separation_vector(first_position, other_position)
                    closest = [entry + shortest[i] for i, entry in enumerate(first_position)]
                    center_position = [entry + closest[i] * node.children[k].weight
                                       for i, entry in enumerate(center_position)]
                setting.periodic_boundaries.correct_position(center_position)
                node.value = Particle(position=center_position)
            return nodes

    tmp = tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False)
    tmp.write(FACTOR_FILE)
    tmp.close()
    maps = FactorTypeMaps(tmp.name)
    cells = CuboidPeriodicCells(cells_per_side=[system["cells_per_side"]], neighbor_layers=1)
    occupancy = SingleActiveCellOccupancy(cells, cell_level=2, maximum_number_occupants=-1)
    pair = ["nearby_sphere", "dipole", "cell_boundary"]
    eoc = ["end_of_chain"] + pair
    taggers = [
        ExcludedCellsTagger(create=pair, trash=pair,
                            event_handler=TwoLeafUnitEventHandler(HardSpherePotential(radius=system["radius"])),
                            number_event_handlers=40, internal_state_label="single_active_cell_occupancy",
                            tag="nearby_sphere"),
        FactorTypeMapInStateTagger(create=pair, trash=pair, event_handler=TwoLeafUnitEventHandler(
            HardDipolePotential(minimum_separation=system["minimum_separation"],
                                maximum_separation=system["maximum_separation"])),
            number_event_handlers=1, factor_type_maps=maps, tag="dipole"),
        CellBoundaryTagger(create=pair, trash=pair, event_handler=CellBoundaryEventHandler(),
                           internal_state_label="single_active_cell_occupancy", tag="cell_boundary"),
        NoInStateTagger(create=["polarization_sampling"], trash=["polarization_sampling"],
                        tag="polarization_sampling",
                        event_handler=FixedIntervalSamplingEventHandler(
                            sampling_interval=system["sampling_interval"], output_handler="dummy_output_handler",
                            first_event_time_zero=True)),
        ActiveGlobalStateInStateTagger(
            create=eoc, trash=eoc, tag="end_of_chain",
            event_handler=SingleIndependentActivePeriodicDirectionEndOfChainEventHandler(
                chain_time=system["chain_time"])),
        NoInStateTagger(create=["end_of_run"], trash=eoc + ["polarization_sampling", "start_of_run", "end_of_run"],
                        tag="end_of_run",
                        event_handler=FinalTimeEndOfRunEventHandler(end_of_run_time=system["end_time"])),
        NoInStateTagger(create=pair + ["polarization_sampling", "end_of_chain", "end_of_run"], trash=["start_of_run"],
                        tag="start_of_run", event_handler=InitialChainStartOfRunEventHandler(
                            initial_direction_of_motion=0, speed=1.0, initial_active_identifier=[0, 0])),
    ]
    tag_index = {t.tag: i for i, t in enumerate(taggers)}
    activator = TagActivator(taggers, internal_states=[occupancy])
    scheduler = ListScheduler()
    state_handler = TreeStateHandler(TreePhysicalState(), TreeLiftingState())
    io = InputOutputHandler(PdbTree(), [DummyOutputHandler()])
    random.seed(system["seed"])
    mediator = SingleProcessMediator(io, state_handler, scheduler, activator)
    os.unlink(tmp.name)

    def unit(identifier):
        return identifier[0] * P + identifier[1]

    def leaf_positions():
        return np.array([state_handler._physical_state.get((r, l)).value.position for r in range(R) for l in range(P)])

    initial = leaf_positions()
    charges = np.array([state_handler._physical_state.get((r, l)).value.charge["electric"]
                        for r in range(R) for l in range(P)])
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
        active = [k for k in state_handler._lifting_state._lifting_dictionary.keys() if len(k) == 2]
        pending["eh"] = eh
        pending["active"] = unit(active[0]) if active else -1
        return eh

    depth = [0]

    def insert(out_state):
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
            target = unit(eh._leaf_units[eh._active_leaf_unit_index ^ 1].identifier)
        elif isinstance(eh, SingleIndependentActivePeriodicDirectionEndOfChainEventHandler):
            target = unit(mediator._out_state_arguments[eh][0][0])
        a = pending["active"]
        pos = [0.0, 0.0, 0.0]
        if a >= 0:
            p = state_handler._physical_state.get((a // P, a % P)).value.position
            pos[:len(p)] = p
        new_active = [k for k in state_handler._lifting_state._lifting_dictionary.keys() if len(k) == 2]
        record["tagger"].append(tag_index[tag])
        record["active"].append(a)
        record["target"].append(target)
        record["tq"].append(t.quotient)
        record["tr"].append(t.remainder)
        record["pos"].append(pos)
        record["new_active"].append(unit(new_active[0]) if new_active else -1)
        if len(record["tagger"]) >= system["events"]:
            raise _Stop()

    scheduler.push_event = push_event
    scheduler.get_succeeding_event = get_succeeding_event
    state_handler.insert_into_global_state = insert
    try:
        mediator.run()
    except _Stop:
        pass
    final = leaf_positions()
    return {k: np.asarray(v) for k, v in record.items()}, initial, charges, final, [t.tag for t in taggers]


def main():
    system = SYSTEM
    record, initial, charges, final, tags = run_reference(system)
    meta = dict(system)
    meta["taggers"] = tags
    with open(os.path.join(OUT, system["name"] + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    np.savez_compressed(os.path.join(OUT, system["name"] + ".npz"), positions=initial, charges=charges,
                        final_positions=final, **record)
    print(system["name"], len(record["tagger"]), "events", np.bincount(record["tagger"]))


if __name__ == "__main__":
    main()
