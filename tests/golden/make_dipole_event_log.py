"""Write the event-log fixture of a composite-object run (dipoles) by running the reference
JeLLyFysh (pure-Python hot path) under CPython in this container.

Why: nothing in the reference's own tests pins the event sequence of composite point objects:
the factor-type-map tagger's candidate order (factor_type_maps.py:325-356,
factor_type_map_in_state_tagger.py:83-107), the leaf-level end-of-chain draws
(single_independent_active_periodic_direction_end_of_chain_event_handler.py:213-237) and the
dipole placement of the random input handler (dipole_random_node_creator.py:69-99).  This
fixture does.  The reference never travels with the repository: only the resulting
inputs/outputs (.npz of plain arrays + .json parameters) are committed.

The run: 6 dipoles (two leaf units of charge +1 / -1 each) placed by the reference's
RandomInputHandler with its DipoleRandomNodeCreator from random.seed(seed), in the unit cube at
beta = 1, with three factor types of the FactorTypeMapInStateTagger:
  Harmonic   [0, 1]                      DisplacedEvenPowerPotential(r0 = 0.1, power 2, k = 200)
  Repulsive  [0, 3], [1, 2]              InversePowerPotential(power 6, prefactor 1e-6)
  Soft       [0, 2], [0, 3], [1, 2], [1, 3]  InversePowerPotential(power 12, prefactor 1e-9)
(the first two as in config_files/2018_JCP_149_064113/dipoles/atom_factors.ini, whose Coulomb
factor needs the reference's cffi-compiled module, absent here), sampling every 0.56789, chains of
0.78965 cycling +x, +y, +z, list scheduler, no cells.

Usage:  python tests/golden/make_dipole_event_log.py [/path/to/reference]

Adjustments, all outside the reference's algorithm:
  * FactorTypeMapInStateTagger.yield_identifiers_send_event_time yields its factors through a
    ``set``: CPython iterates it in hash order, PyPy (the reference's recommended interpreter)
    in insertion order.  The fixture takes the insertion order (``dict.fromkeys``).
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

FACTOR_FILE = """[0, 1], Harmonic
[0, 3], Repulsive
[1, 2], Repulsive
[0, 2], Soft
[0, 3], Soft
[1, 2], Soft
[1, 3], Soft
"""

SYSTEM = dict(name="dipoles_factor_maps_r6", dim=3, roots=6, leaves=2, system_length=1.0, beta=1.0, seed=17,
              chain_time=0.78965, sampling_interval=0.56789, end_time=1.0e9, scheduler="list", events=12000,
              max_dipole_separation=0.2, charges=[1.0, -1.0],
              factors=[dict(tag="harmonic", kind="displaced_even_power", equilibrium_separation=0.1, power=2,
                            prefactor=200.0, local=True, targets=[[1], [0]]),
                       dict(tag="repulsive", kind="inverse_power", power=6.0, prefactor=1e-6, local=False,
                            targets=[[3], [2]]),
                       dict(tag="soft", kind="inverse_power", power=12.0, prefactor=1e-9, local=False,
                            targets=[[2, 3], [2, 3]])])


class _Stop(Exception):
    pass


def run_reference(system):
    for name in list(sys.modules):
        if name == "jellyfysh" or name.startswith("jellyfysh."):
            del sys.modules[name]
    sys.dont_write_bytecode = True
    sys.path.insert(0, REF)
    try:
        import jellyfysh.setting as setting
        from jellyfysh.setting.hypercubic_setting import HypercubicSetting
        from jellyfysh.activator.tag_activator import TagActivator
        from jellyfysh.activator.tagger.factor_type_maps import FactorTypeMaps
        from jellyfysh.activator.tagger.factor_type_map_in_state_tagger import FactorTypeMapInStateTagger
        from jellyfysh.activator.tagger.active_global_state_in_state_tagger import ActiveGlobalStateInStateTagger
        from jellyfysh.activator.tagger.no_in_state_tagger import NoInStateTagger
        from jellyfysh.base.node import yield_leaf_nodes
        from jellyfysh.event_handler.two_leaf_unit_event_handler import TwoLeafUnitEventHandler
        from jellyfysh.event_handler.single_independent_active_periodic_direction_end_of_chain_event_handler \
            import SingleIndependentActivePeriodicDirectionEndOfChainEventHandler
        from jellyfysh.event_handler.final_time_end_of_run_event_handler import FinalTimeEndOfRunEventHandler
        from jellyfysh.event_handler.initial_chain_start_of_run_event_handler import \
            InitialChainStartOfRunEventHandler
        from jellyfysh.event_handler.fixed_interval_sampling_event_handler import FixedIntervalSamplingEventHandler
        from jellyfysh.potential.inverse_power_potential import InversePowerPotential
        from jellyfysh.potential.displaced_even_power_potential import DisplacedEvenPowerPotential
        from jellyfysh.scheduler.list_scheduler import ListScheduler
        from jellyfysh.state_handler.tree_state_handler import TreeStateHandler
        from jellyfysh.state_handler.physical_state.tree_physical_state import TreePhysicalState
        from jellyfysh.state_handler.lifting_state.tree_lifting_state import TreeLiftingState
        from jellyfysh.input_output_handler.input_output_handler import InputOutputHandler
        from jellyfysh.input_output_handler.input_handler.random_input_handler import RandomInputHandler
        from jellyfysh.input_output_handler.input_handler.random_node_creator.dipole_random_node_creator import \
            DipoleRandomNodeCreator
        from jellyfysh.input_output_handler.input_handler.charge_values import ChargeValues
        from jellyfysh.input_output_handler.output_handler.dummy_output_handler import DummyOutputHandler
        from jellyfysh.mediator.single_process_mediator import SingleProcessMediator
    finally:
        sys.path.pop(0)

    def ordered_identifiers(self, extracted_active_global_state):
        yield from dict.fromkeys(factor for root_cnode in extracted_active_global_state
                                 for leaf_cnode in yield_leaf_nodes(root_cnode)
                                 for factor in self._factor_type_map.yield_factor_identifier(
                                     leaf_cnode.value.identifier))
    FactorTypeMapInStateTagger.yield_identifiers_send_event_time = ordered_identifiers

    P, R = system["leaves"], system["roots"]
    HypercubicSetting(beta=system["beta"], dimension=system["dim"], system_length=system["system_length"])
    creator = DipoleRandomNodeCreator(min_initial_dipole_separation=0.0,
                                      max_initial_dipole_separation=system["max_dipole_separation"],
                                      charge_values=[ChargeValues(system["charges"], "electric_charge")])
    input_handler = RandomInputHandler(creator, number_of_root_nodes=R)
    tmp = tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False)
    tmp.write(FACTOR_FILE)
    tmp.close()
    maps = FactorTypeMaps(tmp.name)
    pair_tags = [f["tag"] for f in system["factors"]]
    eoc = ["end_of_chain"] + pair_tags
    taggers = []
    for f in system["factors"]:
        if f["kind"] == "displaced_even_power":
            pot = DisplacedEvenPowerPotential(equilibrium_separation=f["equilibrium_separation"], power=f["power"],
                                              prefactor=f["prefactor"])
        else:
            pot = InversePowerPotential(power=f["power"], prefactor=f["prefactor"])
        taggers.append(FactorTypeMapInStateTagger(create=pair_tags, trash=pair_tags,
                                                  event_handler=TwoLeafUnitEventHandler(pot),
                                                  number_event_handlers=2 * R * P, factor_type_maps=maps,
                                                  tag=f["tag"]))
    taggers.append(NoInStateTagger(create=["sampling"], trash=["sampling"], tag="sampling",
                                   event_handler=FixedIntervalSamplingEventHandler(
                                       sampling_interval=system["sampling_interval"],
                                       output_handler="dummy_output_handler")))
    taggers.append(ActiveGlobalStateInStateTagger(
        create=eoc, trash=eoc, tag="end_of_chain",
        event_handler=SingleIndependentActivePeriodicDirectionEndOfChainEventHandler(chain_time=system["chain_time"])))
    taggers.append(NoInStateTagger(create=["end_of_run"], trash=eoc + ["sampling", "end_of_run"], tag="end_of_run",
                                   event_handler=FinalTimeEndOfRunEventHandler(end_of_run_time=system["end_time"])))
    taggers.append(NoInStateTagger(create=pair_tags + ["sampling", "end_of_chain", "end_of_run"],
                                   trash=["start_of_run"], tag="start_of_run",
                                   event_handler=InitialChainStartOfRunEventHandler(
                                       initial_direction_of_motion=0, speed=1.0, initial_active_identifier=[0, 0])))
    tag_index = {t.tag: i for i, t in enumerate(taggers)}
    activator = TagActivator(taggers)
    scheduler = ListScheduler()
    state_handler = TreeStateHandler(TreePhysicalState(), TreeLiftingState())
    io = InputOutputHandler(input_handler, [DummyOutputHandler()])
    random.seed(system["seed"])
    mediator = SingleProcessMediator(io, state_handler, scheduler, activator)
    os.unlink(tmp.name)

    def unit(identifier):
        return identifier[0] * P + identifier[1]

    def leaf_positions():
        return np.array([state_handler._physical_state.get((r, l)).value.position for r in range(R) for l in range(P)])

    initial = leaf_positions()
    charges = np.array([state_handler._physical_state.get((r, l)).value.charge["electric_charge"]
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
            pos = list(state_handler._physical_state.get((a // P, a % P)).value.position)
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
