"""Output handlers of a run, as the reference's input_output_handler/output_handler package
defines them, fed from the device chain at its sampling events.

- ``get_uuid`` -- base/uuid.py:20-36: one random run identification hash per process.
- ``HardBufferedTextWriter`` -- output_handler.py:90-128: writes to ``filename + '.tmp'``, starting
  with the line ``# Run identification hash: <uuid>``, and renames the file to ``filename`` when
  closed.
- ``OutputHandler`` -- output_handler.py:30-88: checks that the output directory exists and
  reports every hundredth sample.
- ``SeparationOutputHandler`` -- separation_output_handler.py:33-99 for point masses: at every
  sample, the norms of the shortest separations of all pairs i < j, one per line as Python
  prints floats.  The norms are computed on the device (``ecmc_separation_norms``).
- ``DummyOutputHandler`` -- dummy_output_handler.py: takes samples and writes nothing.

The names are the INI aliases the reference uses (``separation_output_handler``,
``dummy_output_handler``); ``make_output_handler`` builds one from its INI section.
"""
import ctypes
import os
import uuid

import numpy as np

from . import _lib as L

_UUID = None


class ConfigurationError(ValueError):
    """base.exceptions.ConfigurationError of the reference."""


def get_uuid():
    """The run identification hash: generated on the first call, returned on every call."""
    global _UUID
    if _UUID is None:
        _UUID = uuid.uuid4()
    return _UUID


class HardBufferedTextWriter:
    """A file-like ``write`` target for ``print``; the text goes to ``filename + '.tmp'`` and the
    file takes its name on ``close``."""

    def __init__(self, filename):
        self._filename = filename
        self._tmp_file = open(filename + ".tmp", "w")
        print("# Run identification hash: {0}".format(get_uuid()), file=self)

    def write(self, string):
        self._tmp_file.write(string)

    def close(self):
        self._tmp_file.close()
        os.replace(self._filename + ".tmp", self._filename)


class OutputHandler:
    def __init__(self, output_filename):
        self._output_filename = output_filename
        self._counter = 0
        directory_name = os.path.dirname(output_filename)
        if directory_name and not os.path.isdir(directory_name):
            raise ConfigurationError("Output directory '{0}' does not exist.".format(directory_name))

    def write(self, *args):
        self._counter += 1
        if self._counter % 100 == 0:
            print("{0}: Calculated {1} samples.".format(self.__class__.__name__, self._counter))

    def post_run(self):
        raise NotImplementedError


class DummyOutputHandler(OutputHandler):
    def __init__(self):
        super().__init__(".")

    def write(self, *args):
        pass

    def post_run(self):
        pass


class SeparationOutputHandler(OutputHandler):
    """Norms of the pair separations at every sample
    (input_output_handler/output_handler/separation_output_handler.py:46-106): for point masses
    every pair i < j into one file; with composite objects of P leaf units, every pair of leaf
    units of different objects (objects i < j, then leaves in order), into P files by the distance
    of their leaf numbers, named <stem>_1<d + P + 1>.<suffix> (for dipoles _13: equal leaf
    numbers, _14: different ones).

    ``write(positions, system_length)`` takes the (n, dim) positions of the leaf units in
    identifier order (unit root * P + leaf)."""

    def __init__(self, filename, leaves_per_root=1):
        super().__init__(filename)
        parts = filename.split(".")
        if len(parts) != 2:
            raise AssertionError("the filename must contain one file format suffix")
        self.leaves_per_root = P = int(leaves_per_root)
        if P > 1:
            self._files = [HardBufferedTextWriter("%s_1%d.%s" % (parts[0], d + P + 1, parts[1])) for d in range(P)]
        else:
            self._files = [HardBufferedTextWriter(filename)]

    @staticmethod
    def separation_norms(positions, system_length):
        """The n (n - 1) / 2 norms, computed on the device, in the handler's order."""
        import torch
        L.require_device()
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        n, dim = pos.shape
        out = torch.empty(n * (n - 1) // 2, dtype=torch.float64, device="cuda")
        if n < 2:
            return out.cpu().numpy()
        d_pos = torch.from_numpy(pos).to("cuda")
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        L.check(L.lib().ecmc_separation_norms(dim, n, L.dptr(d_pos), float(system_length), L.dptr(out), stream))
        torch.cuda.current_stream().synchronize()
        return out.cpu().numpy()

    @staticmethod
    def composite_pairs(n, leaves_per_root):
        """(index into the i < j pair list, leaf-number distance) of every pair of leaf units of
        different composite objects, in the reference's loop order."""
        P = int(leaves_per_root)
        R = n // P
        out = []
        for i in range(R):
            for j in range(i + 1, R):
                for a in range(P):
                    for b in range(P):
                        u, v = i * P + a, j * P + b
                        out.append((u * n - u * (u + 1) // 2 + (v - u - 1), abs(a - b)))
        return out

    def write(self, positions, system_length):
        super().write(positions, system_length)
        norms = self.separation_norms(positions, system_length)
        # print(float) writes repr(float): the shortest string that reads back as the same double
        if self.leaves_per_root == 1:
            self._files[0].write("".join(repr(float(v)) + "\n" for v in norms))
            return
        lines = [[] for _ in self._files]
        for k, d in self.composite_pairs(len(positions), self.leaves_per_root):
            lines[d].append(repr(float(norms[k])) + "\n")
        for f, ls in zip(self._files, lines):
            f.write("".join(ls))

    def post_run(self):
        for f in self._files:
            f.close()


class PolarizationOutputHandler(OutputHandler):
    """input_output_handler/output_handler/polarization_output_handler.py:35-110: at every sample
    the polarization sum over charge-neutral composite objects of charge x position, each leaf
    position taken at the periodic image closest to its object (for two leaves of charges q, -q
    that is q times the shortest separation from the second leaf to the first), one line of
    tab-separated components under a '# Polarization Vector' header.

    ``write(positions, system_length, charges)`` takes the (n, dim) leaf positions in identifier
    order and the charges of the handler's charge."""

    def __init__(self, filename, charge, leaves_per_root=2):
        super().__init__(filename)
        if int(leaves_per_root) < 2:
            raise ConfigurationError("The output handler PolarizationOutputHandler can only be used with charge neutral "
                                     "composite point objects.")
        self.leaves_per_root = int(leaves_per_root)
        self.charge = charge
        self._file = HardBufferedTextWriter(filename)
        print("# Polarization Vector", file=self._file)

    def polarization(self, positions, system_length, charges):
        P = self.leaves_per_root
        pos = np.asarray(positions, dtype=np.float64).reshape(-1, P, np.shape(positions)[-1])
        q = np.asarray(charges, dtype=np.float64).reshape(-1, P)
        if np.any(q.sum(axis=1) != 0.0):
            raise AssertionError("every composite object must be charge neutral")
        L_ = float(system_length)
        # images closest to the first leaf (for charge-neutral objects the reference's closest
        # images to the object's position differ from these by a shift that drops out)
        s = np.mod((pos - pos[:, :1, :]) + L_ / 2, L_) - L_ / 2
        return (q[:, :, None] * s).sum(axis=(0, 1))

    def write(self, positions, system_length, charges):
        super().write(positions, system_length, charges)
        print("\t".join(map(str, (float(v) for v in self.polarization(positions, system_length, charges)))),
              file=self._file)

    def post_run(self):
        self._file.close()


_HANDLERS = {"separation_output_handler": SeparationOutputHandler, "dummy_output_handler": DummyOutputHandler,
             "polarization_output_handler": PolarizationOutputHandler}


def make_output_handler(cls, options, leaves_per_root=1):
    """An output handler from its INI class name (snake case) and section options."""
    if cls not in _HANDLERS:
        raise ConfigurationError("output handler %s is not supported" % cls)
    if cls in ("separation_output_handler", "polarization_output_handler"):
        return _HANDLERS[cls](leaves_per_root=leaves_per_root, **options)
    return _HANDLERS[cls](**options)
