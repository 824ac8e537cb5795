"""Input handlers: the initial global physical state of a run, as the reference's
input_output_handler/input_handler package builds it.

- ``ChargeValues`` -- charge_values.py:25-90: one charge per child of a root node, and a name.
- ``RandomInputHandler`` with ``AtomRandomNodeCreator`` -- random_input_handler.py:34-92,
  atom_random_node_creator.py:26-60, hypercubic_setting.py:226-235: every root node is a point
  mass at ``[random.uniform(0.0, L) for _ in range(dimension)]``, drawn from the run's random
  stream (CPython's ``random``, the reference's generator).  ``read`` also reports how many
  words of the stream the draws took (two per ``uniform``), so that the device chain starts
  its own draws after them (``rng_skip``).
- ``DipoleRandomNodeCreator`` -- dipole_random_node_creator.py:69-99 with
  vectors.py:242-262: a random centre, a direction by rejection (components uniform in
  [-1, 1] until 0 < norm <= 1, then divided by the norm), a half-separation uniform in
  [min / 2, max / 2], the two atoms at centre +- direction x half-separation, each coordinate
  taken % L; the first atom gets each charge's first value, the second its second.
- ``PdbInputHandler`` -- pdb_input_handler.py:33-191.  The reference reads the file through
  MDAnalysis (not installed here); this is a restatement of what that reader yields for the
  fixed-column PDB records the reference writes and tests: the box from CRYST1 (columns 7-33),
  each ATOM/HETATM's serial (7-11), residue number (23-26) and coordinates (31-54) stored as
  float32, as MDAnalysis stores them.  Positions are corrected into [0, L) with Python's ``%``
  (hypercubic_setting.py:101-117), and the reference's checks on the file are kept.

Only point masses run on the device chain; composite objects (residues of several atoms) are
read, with their atoms grouped per residue, but a run with them raises ConfigurationError.
"""
import os

import numpy as np

from .output import ConfigurationError


class ChargeValues:
    def __init__(self, charge_values, charge_name=None):
        self.charge_values = [float(v) for v in charge_values]
        self.charge_name = charge_name if charge_name is not None else "charge_values"

    def __len__(self):
        return len(self.charge_values)

    def __getitem__(self, item):
        return self.charge_values[item]


class AtomRandomNodeCreator:
    number_of_nodes_per_root_node = 1
    number_of_node_levels = 1

    def __init__(self, charge_values=()):
        self.charge_values = list(charge_values)

    def fill_root_node(self, rng, dimension, system_length):
        position = [rng.uniform(0.0, system_length) for _ in range(dimension)]
        return position, {c.charge_name: c[0] for c in self.charge_values}


class DipoleRandomNodeCreator:
    number_of_nodes_per_root_node = 2
    number_of_node_levels = 2

    def __init__(self, min_initial_dipole_separation=0.0, max_initial_dipole_separation=0.05, charge_values=()):
        self.min_initial_dipole_separation = float(min_initial_dipole_separation)
        self.max_initial_dipole_separation = float(max_initial_dipole_separation)
        self.charge_values = list(charge_values)

    @staticmethod
    def random_vector_on_unit_sphere(rng, dimension):
        while True:
            vector = [rng.uniform(-1, 1) for _ in range(dimension)]
            vector_norm = sum(c * c for c in vector) ** 0.5
            if 0.0 < vector_norm <= 1.0:
                break
        return [v / vector_norm * 1.0 for v in vector]

    def fill_root_node(self, rng, dimension, system_length):
        """([atom one, atom two] positions, {name: [charge one, charge two]})."""
        center = [rng.uniform(0.0, system_length) for _ in range(dimension)]
        direction = self.random_vector_on_unit_sphere(rng, dimension)
        half = rng.uniform(self.min_initial_dipole_separation / 2.0, self.max_initial_dipole_separation / 2.0)
        one = [_py_mod(center[d] + direction[d] * half, system_length) for d in range(dimension)]
        two = [_py_mod(center[d] - direction[d] * half, system_length) for d in range(dimension)]
        return [one, two], {c.charge_name: [c[0], c[1]] for c in self.charge_values}


class _CountingRandom:
    """The run's random.Random with a count of the 32-bit words its uniform draws take
    (random() combines two)."""

    def __init__(self, rng):
        self._rng = rng
        self.words = 0

    def uniform(self, a, b):
        self.words += 2
        return self._rng.uniform(a, b)


class RandomInputHandler:
    def __init__(self, number_of_root_nodes, random_node_creator):
        self.number_of_root_nodes = int(number_of_root_nodes)
        self.random_node_creator = random_node_creator

    def read(self, rng, dimension, system_length):
        """Positions (n, dim) -- (n, children, dim) for composite objects --, charges {name: (n,)
        or (n, children)} and the number of random words drawn."""
        counting = _CountingRandom(rng)
        positions, charges = [], {}
        for _ in range(self.number_of_root_nodes):
            p, c = self.random_node_creator.fill_root_node(counting, dimension, system_length)
            positions.append(p)
            for k, v in c.items():
                charges.setdefault(k, []).append(v)
        return (np.array(positions, dtype=np.float64), {k: np.array(v) for k, v in charges.items()},
                counting.words)


def _py_mod(x, y):
    return x % y


class PdbInputHandler:
    def __init__(self, filename, charge_values=(), dimension=3, system_length=1.0):
        if not filename.endswith(".pdb"):
            raise ConfigurationError("Input filename for input handler PdbInputHandler should end with .pdb.")
        if dimension > 3:
            raise ConfigurationError("The PDB format does not allow for dimensions > 3.")
        if not os.path.isfile(filename):
            raise OSError("The file {0} does not exist.".format(filename))
        self.dimension = dimension
        self.system_length = system_length
        self.charge_values = list(charge_values)
        self.box, self.atoms = self._parse(filename)
        residues = {}
        for a in self.atoms:
            residues.setdefault(a["resid"], []).append(a)
        self.residues = [residues[k] for k in sorted(residues)]
        sizes = {len(r) for r in self.residues}
        if len(sizes) != 1:
            raise ConfigurationError("Currently only point composite objects with the same number of children are "
                                     "supported!")
        self.number_of_root_nodes = len(self.residues)
        self.number_of_nodes_per_root_node = sizes.pop()
        self.number_of_node_levels = 1 if self.number_of_nodes_per_root_node == 1 else 2
        for index, length in enumerate(self.box[:dimension]):
            if abs(length - system_length) > 1.0e-5:
                raise ConfigurationError("The system lengths {0} stored in the .pdb file {1} are not equal to the "
                                         "system lengths {2}.".format(list(self.box[:dimension]), filename,
                                                                      [system_length] * dimension))
        if any(v != 0.0 for v in self.box[dimension:3]):
            raise ConfigurationError("The dimension {0} < 3 but the superfluous system lengths {1} stored in the .pdb "
                                     "file {2} are not 0.0.".format(dimension, list(self.box[dimension:3]), filename))
        for c in self.charge_values:
            if len(c) != self.number_of_nodes_per_root_node:
                raise ConfigurationError("Please give the charge {0} for every child of a composite point object!"
                                         .format(c.charge_name))
        if len({c.charge_name for c in self.charge_values}) != len(self.charge_values):
            raise ConfigurationError("Each given charge should have a different name!")

    @staticmethod
    def _parse(filename):
        box = np.zeros(3, dtype=np.float32)
        atoms = []
        with open(filename) as f:
            for line in f:
                rec = line[:6].strip()
                if rec == "CRYST1":
                    box = np.array([float(line[6:15]), float(line[15:24]), float(line[24:33])], dtype=np.float32)
                elif rec in ("ATOM", "HETATM"):
                    xyz = np.array([float(line[30:38]), float(line[38:46]), float(line[46:54])], dtype=np.float32)
                    atoms.append(dict(id=int(line[6:11]), resid=int(line[22:26]), position=xyz))
                elif rec == "END":
                    break
        return [float(v) for v in box], atoms

    def read(self):
        """Per residue (root node), the corrected positions of its atoms in file order, and per
        charge name the charges, as (root nodes, atoms per root node) arrays."""
        n, k = self.number_of_root_nodes, self.number_of_nodes_per_root_node
        pos = np.zeros((n, k, self.dimension))
        charges = {c.charge_name: np.zeros((n, k)) for c in self.charge_values}
        for r, residue in enumerate(self.residues):
            for j, atom in enumerate(residue):
                p = [float(v) for v in atom["position"]]
                assert all(v == 0.0 for v in p[self.dimension:3])
                pos[r, j] = [_py_mod(p[d], self.system_length) for d in range(self.dimension)]
                atom_index = (atom["id"] - 1) % k
                for c in self.charge_values:
                    charges[c.charge_name][r, j] = c[atom_index]
        return pos, charges

    def read_point_masses(self):
        """(n, dim) positions and {name: (n,)} charges of a file of point masses."""
        if self.number_of_node_levels != 1:
            raise ConfigurationError("composite objects are not run by the device chain")
        pos, charges = self.read()
        return pos[:, 0, :], {k: v[:, 0] for k, v in charges.items()}
