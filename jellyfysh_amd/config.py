"""INI front end: the reference's configuration files drive the device chain.

``RunConfig.from_file(path)`` reads a JeLLyFysh INI file with the reference's factory
conventions (base/factory.py:41-151): keys and values in snake_case, sections in
CamelCase, ``alias (class_name)`` values whose options sit in section [Alias], lists as
comma-separated values.  It accepts the single-process run
(mediator/single_process_mediator.py) of a hypercubic setting with the TagActivator, the
SingleActiveCellOccupancy over CuboidPeriodicCells, the heap or list scheduler and these
taggers / event handlers / potentials:

  taggers        excluded_cells_tagger, factor_type_map_in_state_tagger (point masses: every
                 pair; composite objects: the pairs of its factor type map file),
                 surplus_cells_tagger, cell_bounding_potential_tagger, cell_veto_tagger,
                 cell_boundary_tagger, no_in_state_tagger, active_global_state_in_state_tagger
  handlers       two_leaf_unit_event_handler, two_leaf_unit_bounding_potential_event_handler,
                 two_leaf_unit_cell_bounding_potential_event_handler (cell_bounding_potential
                 with an inner_point_estimator), leaf_unit_cell_veto_event_handler (with an
                 inner_point_estimator), cell_boundary_event_handler,
                 fixed_interval_sampling_event_handler,
                 single_independent_active_periodic_direction_end_of_chain_event_handler,
                 final_time_end_of_run_event_handler, initial_chain_start_of_run_event_handler
  potentials     hard_sphere_potential, inverse_power_potential, lennard_jones_potential,
                 displaced_even_power_potential, merged_image_coulomb_potential,
                 inverse_power_coulomb_bounding_potential, and truncated_inverse_power_potential
                 (no reference counterpart: an inverse power truncated and shifted at a cutoff)

Anything else raises ConfigurationError, as the reference's factory does for options a
class does not take.  The initial configuration comes from the file's input handler (random
placement of point masses or dipoles from the run's random stream, or a .pdb file) or is passed
in by the caller; ``seed`` is the seed of the random stream.  Composite point objects (the
dipole random node creator) run as leaf units, unit root * leaves + leaf, with their
factor_type_map_in_state_tagger's factor type map read from its file (two-leaf factors).

``SingleProcessMediator(config, positions, charges=None, seed=0).run(on_sample=None)``
then runs the whole leg loop on the device until the end-of-run event; at each sampling
event (fixed_interval_sampling_event_handler) it hands the positions to ``on_sample``.
"""
import configparser
import os
import re

import numpy as np

from .chain import DeviceChain, make_desc
from .input_handlers import (AtomRandomNodeCreator, ChargeValues, DipoleRandomNodeCreator, PdbInputHandler,
                             RandomInputHandler)
from .output import ConfigurationError, make_output_handler


def camel(snake):
    return "".join(w[:1].upper() + w[1:] for w in snake.split("_"))


def split_alias(value):
    """'alias (class_name)' -> (alias, class_name); 'class_name' -> (class_name, class_name)."""
    m = re.fullmatch(r"\s*([A-Za-z0-9_]+)\s*(?:\(\s*([A-Za-z0-9_]+)\s*\))?\s*", value)
    if not m:
        raise ConfigurationError("cannot read %r as 'alias (class_name)'" % value)
    return m.group(1), m.group(2) or m.group(1)


def split_list(value):
    """Comma-separated list; commas inside parentheses do not split."""
    out, depth, cur = [], 0, ""
    for ch in value:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


# potential class -> (device potential, constructor arguments in the C ABI's par order,
# their defaults)
_POTENTIALS = {
    "hard_sphere_potential": ("hard_sphere", [("radius", None)]),
    "hard_dipole_potential": ("hard_dipole", [("minimum_separation", None), ("maximum_separation", None)]),
    "inverse_power_potential": ("inverse_power", [("power", None), ("prefactor", None)]),
    "truncated_inverse_power_potential": ("inverse_power_truncated",
                                          [("power", None), ("prefactor", None), ("cutoff", None)]),
    "lennard_jones_potential": ("lennard_jones", [("prefactor", 1.0), ("characteristic_length", 1.0)]),
    "displaced_even_power_potential": ("displaced_even_power",
                                       [("equilibrium_separation", None), ("power", None), ("prefactor", 1.0)]),
    "merged_image_coulomb_potential": ("merged_image_coulomb",
                                       [("alpha", 3.45), ("fourier_cutoff", 6), ("position_cutoff", 2),
                                        ("prefactor", 1.0)]),
    "inverse_power_coulomb_bounding_potential": ("inverse_power_coulomb_bounding", [("prefactor", 1.5837)]),
}

_HANDLER_OPTIONS = {
    "two_leaf_unit_event_handler": {"potential"},
    "two_leaf_unit_bounding_potential_event_handler": {"potential", "bounding_potential", "charge"},
    "two_leaf_unit_cell_bounding_potential_event_handler": {"potential", "bounding_potential", "charge"},
    "leaf_unit_cell_veto_event_handler": {"estimator", "potential", "charge"},
    "cell_boundary_event_handler": set(),
    "fixed_interval_sampling_event_handler": {"sampling_interval", "output_handler", "first_event_time_zero"},
    "single_independent_active_periodic_direction_end_of_chain_event_handler": {"chain_time"},
    "final_time_end_of_run_event_handler": {"end_of_run_time"},
    "initial_chain_start_of_run_event_handler": {"initial_direction_of_motion", "speed",
                                                 "initial_active_identifier"},
}

_TAGGER_OPTIONS = {"create", "trash", "event_handler", "number_event_handlers", "internal_state_label",
                   "factor_type_maps", "factor_type_maps_label", "tag"}


def read_factor_type_maps(path, leaves_per_root):
    """activator/tagger/factor_type_maps.py:160-213: lines '[i, j, ...], FactorName' (single-line
    comments start with '#'); leaf numbers below leaves_per_root belong to one composite object,
    the others to a second one.  A factor is local when all its numbers lie in one object.  Returns
    {FactorName: {"local": bool, "targets": per leaf number k, the other leaf numbers of the
    two-leaf factors holding k, in file order}}."""
    P = int(leaves_per_root)
    line_re = re.compile(r"(\[(?:[0-9]+, )*[0-9]+\]), ((?:[A-Z][a-z]*)+)")
    out = {}
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                continue
            m = line_re.match(line)
            if m is None:
                raise ConfigurationError("Line %r does not match pattern '[index, index, ...], Factor'" % line)
            idx = [int(v) for v in re.findall(r"[0-9]+", m.group(1))]
            if any(i >= 2 * P for i in idx):
                raise ConfigurationError("point mass index too large for two composite objects of %d leaves" % P)
            local = all(i < P for i in idx)
            entry = out.setdefault(m.group(2), {"local": local, "targets": [[] for _ in range(P)]})
            if entry["local"] != local:
                raise ConfigurationError("factor %s mixes local and non-local lines" % m.group(2))
            if len(idx) != 2 or idx[0] == idx[1]:
                raise ConfigurationError("only factors of two leaf units run on the device (%r)" % line.strip())
            for k in idx:
                if k < P:
                    entry["targets"][k].append(idx[1] if idx[0] == k else idx[0])
    return out


def all_leaf_factor_type_map(leaves_per_root):
    """factor_type_maps.py:359-394: every leaf of every other composite object."""
    P = int(leaves_per_root)
    return {"local": False, "targets": [list(range(P, 2 * P)) for _ in range(P)]}


class RunConfig:
    """The parts of a reference INI file that the device chain runs."""

    def __init__(self, parser):
        self.cp = parser
        run = self._section("Run", {"mediator", "setting"})
        mediator = split_alias(run.get("mediator", "single_process_mediator"))[1]
        if mediator != "single_process_mediator":
            raise ConfigurationError("only the single-process mediator runs on the device (got %s)" % mediator)
        setting = split_alias(run.get("setting", "hypercubic_setting"))[1]
        if setting != "hypercubic_setting":
            raise ConfigurationError("only the hypercubic setting is supported (got %s)" % setting)
        s = self._section("HypercubicSetting", {"system_length", "beta", "dimension"})
        self.system_length = float(s["system_length"])
        self.beta = float(s["beta"])
        self.dim = int(s["dimension"])
        m = self._section("SingleProcessMediator", {"state_handler", "scheduler", "activator",
                                                    "input_output_handler"})
        sched = split_alias(m.get("scheduler", "heap_scheduler"))[1]
        if sched not in ("heap_scheduler", "list_scheduler"):
            raise ConfigurationError("unknown scheduler %s" % sched)
        self.scheduler = "heap" if sched == "heap_scheduler" else "list"
        # output handlers (input_output_handler/input_output_handler.py:43-61): alias -> (class, options)
        self.output_handlers = {}
        self.input_handler = None
        self.leaves_per_root = 1
        self._path = getattr(parser, "_ecmc_path", None)
        if "input_output_handler" in m:
            io = self._section(camel(split_alias(m["input_output_handler"])[0]), {"output_handlers", "input_handler"})
            for item in split_list(io.get("output_handlers", "")):
                alias, cls = split_alias(item)
                opts = dict(self.cp[camel(alias)]) if camel(alias) in self.cp else {}
                self.output_handlers[alias] = (cls, opts)
            if io.get("input_handler"):
                self.input_handler = self._input_handler(io["input_handler"])
                creator = getattr(self.input_handler, "random_node_creator", self.input_handler)
                self.leaves_per_root = int(getattr(creator, "number_of_nodes_per_root_node", 1))
        act = split_alias(m.get("activator", "tag_activator"))[1]
        if act != "tag_activator":
            raise ConfigurationError("only the tag activator is supported (got %s)" % act)
        ta = self._section("TagActivator", {"taggers", "internal_states"})
        self.cells_per_side = None
        self.neighbor_layers = 1
        self.max_occupants = 1
        for item in split_list(ta.get("internal_states", "")):
            alias, cls = split_alias(item)
            if cls != "single_active_cell_occupancy":
                raise ConfigurationError("unsupported internal state %s" % cls)
            occ = self._section(camel(alias), {"cells", "cell_level", "maximum_number_occupants"})
            self.max_occupants = int(occ.get("maximum_number_occupants", 1))
            calias, ccls = split_alias(occ.get("cells", "cuboid_periodic_cells"))
            if ccls != "cuboid_periodic_cells":
                raise ConfigurationError("unsupported cells %s" % ccls)
            cs = self._section(camel(calias), {"cells_per_side", "neighbor_layers"})
            cps = [int(v) for v in split_list(cs["cells_per_side"])]
            self.cells_per_side = cps[0] if len(cps) == 1 else cps
            self.neighbor_layers = int(cs.get("neighbor_layers", 1))
        self.taggers = []
        for item in split_list(ta["taggers"]):
            label, cls = split_alias(item)
            self.taggers.append(self._tagger(label, cls))
        labels = [t["tag"] for t in self.taggers]
        for t in self.taggers:
            for lst in ("create", "trash"):
                for tag in t[lst]:
                    if tag not in labels:
                        raise ConfigurationError("tag %s of tagger %s is no tagger" % (tag, t["tag"]))
        if any(t["kind"] == "pair_cells" for t in self.taggers) and self.cells_per_side is None:
            raise ConfigurationError("excluded_cells_tagger needs a cell occupancy")

    def _charge_values(self, value):
        out = []
        for item in split_list(value or ""):
            alias, cls = split_alias(item)
            if cls != "charge_values":
                raise ConfigurationError("unsupported charge values %s" % cls)
            c = self._section(camel(alias), {"charge_name", "charge_values"})
            out.append(ChargeValues([float(v) for v in split_list(c["charge_values"])], c.get("charge_name", alias)))
        return out

    def _input_handler(self, value):
        """input_output_handler/input_handler: random placement of point masses or a .pdb file."""
        alias, cls = split_alias(value)
        if cls == "random_input_handler":
            h = self._section(camel(alias), {"random_node_creator", "number_of_root_nodes"})
            calias, ccls = split_alias(h["random_node_creator"])
            if ccls == "atom_random_node_creator":
                c = self._section(camel(calias), {"charge_values"})
                creator = AtomRandomNodeCreator(self._charge_values(c.get("charge_values")))
            elif ccls == "dipole_random_node_creator":
                c = self._section(camel(calias), {"charge_values", "min_initial_dipole_separation",
                                                  "max_initial_dipole_separation"})
                creator = DipoleRandomNodeCreator(float(c.get("min_initial_dipole_separation", 0.0)),
                                                  float(c.get("max_initial_dipole_separation", 0.05)),
                                                  self._charge_values(c.get("charge_values")))
            else:
                raise ConfigurationError("random node creator %s has no device counterpart" % ccls)
            return RandomInputHandler(int(h["number_of_root_nodes"]), creator)
        if cls == "pdb_input_handler":
            h = self._section(camel(alias), {"filename", "charge_values"})
            fname = h["filename"]
            if not os.path.exists(fname):
                # the reference's relative path (config_files/...) resolved against this file's folder
                # or the package's configs folder
                for base in ([os.path.dirname(self._path)] if self._path else []) + \
                        [os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs")]:
                    for cand in (os.path.join(base, fname), os.path.join(base, os.path.basename(fname)),
                                 os.path.join(base, os.path.basename(os.path.dirname(fname)), os.path.basename(fname))):
                        if os.path.exists(cand):
                            fname = cand
                            break
                    if os.path.exists(fname):
                        break
            return PdbInputHandler(fname, self._charge_values(h.get("charge_values")), self.dim, self.system_length)
        raise ConfigurationError("unsupported input handler %s" % cls)

    def initial_state(self, seed=0):
        """Positions, charges (the electric charge, or the first given; None without charges) and the
        number of random words the input handler drew: random.seed(seed), then the handler reads,
        then the run continues the same stream."""
        import random
        if self.input_handler is None:
            raise ConfigurationError("the configuration has no input handler; give the positions")
        if isinstance(self.input_handler, RandomInputHandler):
            pos, charges, words = self.input_handler.read(random.Random(seed), self.dim, self.system_length)
        else:
            pos, charges = self.input_handler.read()
            words = 0
        q = charges.get("electric_charge", next(iter(charges.values()), None)) if charges else None
        # composite objects: leaf units in identifier order, unit root * leaves + leaf
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, self.dim)
        if q is not None:
            q = np.asarray(q, dtype=np.float64).reshape(-1)
        return pos, q, words

    # ---- helpers ---------------------------------------------------------------------------
    def _section(self, name, allowed):
        if name not in self.cp:
            return {}
        sec = dict(self.cp[name])
        extra = set(sec) - set(allowed)
        if extra:
            raise ConfigurationError("section [%s] has options %s that the device path does not take"
                                     % (name, sorted(extra)))
        return sec

    def _potential(self, value):
        alias, cls = split_alias(value)
        if cls not in _POTENTIALS:
            raise ConfigurationError("potential %s has no device implementation" % cls)
        name, args = _POTENTIALS[cls]
        sec = self._section(camel(alias), {a for a, _ in args})
        params = []
        for a, default in args:
            if a in sec:
                params.append(float(sec[a]))
            elif default is None:
                raise ConfigurationError("missing argument %s of %s in section [%s]" % (a, cls, camel(alias)))
            else:
                params.append(float(default))
        return name, params

    def _cell_bounding(self, value):
        """cell_bounding_potential (estimator = inner_point_estimator) -> the C ABI's bpar:
        estimator prefactor, points_per_side, target_charge, empirical_bound."""
        alias, cls = split_alias(value)
        if cls != "cell_bounding_potential":
            raise ConfigurationError("the cell-bounding handler needs cell_bounding_potential (got %s)" % cls)
        sec = self._section(camel(alias), {"estimator"})
        return self._estimator(sec.get("estimator", "inner_point_estimator"))[1]

    def _estimator(self, value):
        """inner_point_estimator -> (its potential option or None, [prefactor, points_per_side,
        target_charge, empirical_bound]) with the defaults of inner_point_estimator.py:45-46."""
        ealias, ecls = split_alias(value)
        if ecls != "inner_point_estimator":
            raise ConfigurationError("only the inner_point_estimator is supported (got %s)" % ecls)
        est = self._section(camel(ealias), {"prefactor", "potential", "target_charge", "points_per_side",
                                            "empirical_bound", "periodic_boundaries"})
        if est.get("periodic_boundaries", "true").lower() not in ("1", "yes", "true", "on"):
            raise ConfigurationError("the estimator corrects separations for periodic boundaries here")
        return est.get("potential"), [float(est.get("prefactor", 1.5)), float(est.get("points_per_side", 10)),
                                      float(est.get("target_charge", 1.0)), float(est.get("empirical_bound", "inf"))]

    def _tagger(self, label, cls):
        sec = self._section(camel(label), _TAGGER_OPTIONS)
        t = {"tag": label, "create": split_list(sec.get("create", "")), "trash": split_list(sec.get("trash", ""))}
        if "number_event_handlers" in sec:
            t["number_event_handlers"] = int(sec["number_event_handlers"])
        halias, hcls = split_alias(sec["event_handler"])
        if hcls not in _HANDLER_OPTIONS:
            raise ConfigurationError("event handler %s has no device implementation" % hcls)
        h = self._section(camel(halias), _HANDLER_OPTIONS[hcls])
        if hcls == "two_leaf_unit_cell_bounding_potential_event_handler":
            if cls != "cell_bounding_potential_tagger":
                raise ConfigurationError("the cell-bounding handler runs under cell_bounding_potential_tagger")
            t["kind"] = "cell_bounding"
            t["potential"], t["params"] = self._potential(h["potential"])
            t["bounding_params"] = self._cell_bounding(h["bounding_potential"])
            t["charge"] = "charge" in h
        elif hcls == "leaf_unit_cell_veto_event_handler":
            if cls != "cell_veto_tagger":
                raise ConfigurationError("the cell-veto handler runs under cell_veto_tagger")
            t["kind"] = "cell_veto"
            epot, t["bounding_params"] = self._estimator(h["estimator"])
            # leaf_unit_cell_veto_event_handler.py:81: the estimator's potential unless one is given
            pot = h.get("potential", epot)
            if pot is None:
                raise ConfigurationError("the cell-veto handler needs a potential (its own or the estimator's)")
            t["potential"], t["params"] = self._potential(pot)
            if t["potential"] != "merged_image_coulomb":
                raise ConfigurationError("the cell veto runs with merged_image_coulomb_potential here")
            t["charge"] = "charge" in h
        elif hcls in ("two_leaf_unit_event_handler", "two_leaf_unit_bounding_potential_event_handler"):
            if cls == "excluded_cells_tagger":
                t["kind"] = "pair_cells"
            elif cls == "factor_type_map_in_state_tagger" and self.leaves_per_root > 1:
                t["kind"] = "factor_map"
                t["factor_map"] = self._factor_type_map(sec, label)
            elif cls == "factor_type_map_in_state_tagger":
                t["kind"] = "pair_all"
            elif cls == "surplus_cells_tagger":
                t["kind"] = "surplus"
            else:
                raise ConfigurationError("pair handlers run under excluded_cells_tagger, surplus_cells_tagger or "
                                         "factor_type_map_in_state_tagger (got %s)" % cls)
            t["potential"], t["params"] = self._potential(h["potential"])
            if hcls == "two_leaf_unit_bounding_potential_event_handler":
                t["bounding_potential"], t["bounding_params"] = self._potential(h["bounding_potential"])
                t["charge"] = "charge" in h
        elif hcls == "cell_boundary_event_handler":
            t["kind"] = "cell_boundary"
        elif hcls == "fixed_interval_sampling_event_handler":
            t["kind"] = "sampling"
            t["sampling_interval"] = float(h["sampling_interval"])
            t["first_event_time_zero"] = h.get("first_event_time_zero", "false").lower() in ("1", "yes", "true", "on")
            t["output_handler"] = h.get("output_handler")
        elif hcls == "single_independent_active_periodic_direction_end_of_chain_event_handler":
            t["kind"] = "end_of_chain"
            t["chain_time"] = float(h["chain_time"])
        elif hcls == "final_time_end_of_run_event_handler":
            t["kind"] = "end_of_run"
            t["end_of_run_time"] = float(h["end_of_run_time"])
        else:
            t["kind"] = "start_of_run"
            t["initial_direction_of_motion"] = int(h.get("initial_direction_of_motion", 0))
            t["speed"] = float(h.get("speed", 1.0))
            ident = [int(v) for v in split_list(h.get("initial_active_identifier", "0"))]
            if len(ident) == 2 and self.leaves_per_root > 1:
                ident = [ident[0] * self.leaves_per_root + ident[1]]
            if len(ident) != 1:
                raise ConfigurationError("initial_active_identifier %r does not name a leaf unit"
                                         % h.get("initial_active_identifier"))
            t["initial_active_identifier"] = ident[0]
        return t

    def _factor_type_map(self, sec, label):
        """The factor type map of a factor_type_map_in_state_tagger: its factor_type_maps section's
        file, the entry named by factor_type_maps_label (or the tag) in CamelCase, or every leaf of
        every other composite object when the file has no such factor."""
        import os
        if "factor_type_maps" not in sec:
            raise ConfigurationError("factor_type_map_in_state_tagger needs factor_type_maps")
        alias, cls = split_alias(sec["factor_type_maps"])
        if cls != "factor_type_maps":
            raise ConfigurationError("unsupported factor type maps %s" % cls)
        fname = self._section(camel(alias), {"filename"}).get("filename")
        if not fname:
            raise ConfigurationError("section [%s] needs a filename" % camel(alias))
        here = os.path.dirname(os.path.abspath(__file__))
        candidates = [fname]
        if self._path:
            candidates.append(os.path.join(os.path.dirname(self._path), fname))
        candidates.append(os.path.join(here, "configs", "factor_set_files", os.path.basename(fname)))
        path = next((c for c in candidates if os.path.exists(c)), None)
        if path is None:
            raise ConfigurationError("factor type map file %s not found" % fname)
        maps = read_factor_type_maps(path, self.leaves_per_root)
        name = camel(sec.get("factor_type_maps_label", label))
        return maps.get(name, all_leaf_factor_type_map(self.leaves_per_root))

    # ---- construction ----------------------------------------------------------------------
    @classmethod
    def from_string(cls, text):
        cp = configparser.ConfigParser()
        cp.read_string(text)
        return cls(cp)

    @classmethod
    def from_file(cls, path):
        cp = configparser.ConfigParser()
        with open(path) as f:
            cp.read_file(f)
        cp._ecmc_path = os.path.abspath(path)
        return cls(cp)

    def desc(self, n, seed=0, rng_skip=0):
        cps = self.cells_per_side if self.cells_per_side is not None else 1
        return make_desc(self.dim, n, self.system_length, self.beta, cps, self.taggers, seed,
                         scheduler=self.scheduler, neighbor_layers=self.neighbor_layers,
                         max_occupants=self.max_occupants, rng_skip=rng_skip, leaves_per_root=self.leaves_per_root)


class SingleProcessMediator:
    """mediator/single_process_mediator.py:84-156 with the whole leg loop on the device."""

    def __init__(self, config, positions=None, charges=None, seed=0):
        if isinstance(config, str):
            config = RunConfig.from_file(config)
        self.config = config
        skip = 0
        if positions is None:
            # the configuration's input handler (random placement draws from the run's stream)
            positions, q, skip = config.initial_state(seed)
            if charges is None:
                charges = q
        pos = np.asarray(positions, dtype=np.float64)
        self.charges = None if charges is None else np.asarray(charges, dtype=np.float64)
        # charges reach the device only where a tagger multiplies by them
        uses_charge = any(t.get("charge") for t in config.taggers)
        self.chain = DeviceChain(config.desc(pos.shape[0], seed, rng_skip=skip), pos,
                                 self.charges if uses_charge else None)
        self.sampling = [i for i, t in enumerate(config.taggers) if t["kind"] == "sampling"]
        self.output_handlers = {alias: make_output_handler(cls, opts, config.leaves_per_root)
                                for alias, (cls, opts) in config.output_handlers.items()}
        # sampling tagger -> the output handler its event handler writes to
        self.writers = {i: self.output_handlers[config.taggers[i]["output_handler"]] for i in self.sampling
                        if config.taggers[i].get("output_handler") in self.output_handlers}

    def run(self, on_sample=None, chunk=1 << 20):
        """Run to the end-of-run event; returns the number of committed legs.  At each sampling
        event the positions go to the sampling handler's output handler and to ``on_sample``;
        at the end every output handler's ``post_run`` runs (as the mediator's post_run does)."""
        mask = 0
        for i in self.sampling:
            if on_sample is not None or i in self.writers:
                mask |= 1 << i
        self.chain.set_stop_mask(mask)
        total = 0
        while True:
            if self.writers:
                # a short log tells which tagger's event stopped the run
                ev = self.chain.run(min(chunk, 4096), log=True)
                total += len(ev)
                last = int(ev["tagger"][-1]) if len(ev) else -1
            else:
                total += self.chain.run(chunk, log=False)
                last = -1
            if self.chain.active()[4]:      # end of run committed
                for h in self.output_handlers.values():
                    h.post_run()
                return total
            if mask:
                pos = None
                if last in self.writers:
                    pos = self.chain.positions()
                    w = self.writers[last]
                    if hasattr(w, "polarization"):
                        w.write(pos, self.config.system_length, self.charges)
                    else:
                        w.write(pos, self.config.system_length)
                if on_sample is not None and (last < 0 or last in self.sampling):
                    on_sample(pos if pos is not None else self.chain.positions())


class BatchedRunConfig:
    """An INI file of the batched form (no reference counterpart as a whole; read with the
    reference's factory conventions, base/factory.py:41-151): [Run] mediator =
    batched_replicas_mediator, the [HypercubicSetting]'s beta and dimension 3, and a
    [BatchedReplicasMediator] section with the replica set (number_of_replicas, replica sizes
    log-uniform in [minimum_number_of_units, maximum_number_of_units], densities uniform in
    [minimum_density, maximum_density], lattice starts with a jitter, drawn from
    workload_seed), chain_length, seed (the replicas' Philox key) and one or two potentials,
    each a truncated_inverse_power_potential of power 6 or 12 with one common cutoff.

    ``workload()`` gives (n, system lengths, positions); ``batch_arguments()`` the keyword
    arguments of jellyfysh_amd.batch.DeviceBatch; ``batch()`` builds it."""

    def __init__(self, parser):
        self.cp = parser
        run = RunConfig._section(self, "Run", {"mediator", "setting"})
        if split_alias(run.get("mediator", ""))[1] != "batched_replicas_mediator":
            raise ConfigurationError("not a batched run (mediator = batched_replicas_mediator)")
        s = RunConfig._section(self, "HypercubicSetting", {"beta", "dimension", "system_length"})
        if int(s.get("dimension", 3)) != 3:
            raise ConfigurationError("the batched replicas are three-dimensional")
        if "system_length" in s:
            raise ConfigurationError("each replica's system length follows from its size and density")
        self.beta = float(s["beta"])
        m = RunConfig._section(self, "BatchedReplicasMediator",
                               {"number_of_replicas", "minimum_number_of_units", "maximum_number_of_units",
                                "minimum_density", "maximum_density", "jitter", "workload_seed", "chain_length",
                                "seed", "potentials"})
        self.number_of_replicas = int(m["number_of_replicas"])
        self.n_min, self.n_max = int(m["minimum_number_of_units"]), int(m["maximum_number_of_units"])
        self.rho_min, self.rho_max = float(m["minimum_density"]), float(m["maximum_density"])
        self.jitter = float(m.get("jitter", 0.01))
        self.workload_seed = int(m.get("workload_seed", 0))
        self.chain_length = float(m.get("chain_length", 1.0))
        self.seed = int(m.get("seed", 1))
        self.factors, cutoffs = [], set()
        for item in split_list(m["potentials"]):
            alias, cls = split_alias(item)
            if cls != "truncated_inverse_power_potential":
                raise ConfigurationError("the batched replicas take truncated_inverse_power_potential (got %s)" % cls)
            sec = RunConfig._section(self, camel(alias), {"power", "prefactor", "cutoff"})
            p, k, rc = float(sec["power"]), float(sec["prefactor"]), float(sec["cutoff"])
            if p not in (6.0, 12.0):
                raise ConfigurationError("powers 6 and 12 are supported")
            self.factors.append((p, k))
            cutoffs.add(rc)
        if not 1 <= len(self.factors) <= 2 or len(cutoffs) != 1:
            raise ConfigurationError("one or two factors with one common cutoff")
        self.cutoff = cutoffs.pop()

    @classmethod
    def from_file(cls, path):
        cp = configparser.ConfigParser()
        with open(path) as f:
            cp.read_file(f)
        return cls(cp)

    def workload(self):
        from .batch import batched_lj_workload
        return batched_lj_workload(self.number_of_replicas, self.n_min, self.n_max, self.rho_min, self.rho_max,
                                   seed=self.workload_seed, jitter=self.jitter)

    def batch_arguments(self):
        return dict(beta=self.beta, cutoff=self.cutoff, chain_length=self.chain_length, factors=tuple(self.factors),
                    seed=self.seed)

    def batch(self, classes=None):
        from .batch import DeviceBatch
        n, lengths, pos = self.workload()
        return DeviceBatch(n, lengths, pos, classes=classes, **self.batch_arguments())
