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
    """Norms of all pair separations of point masses at every sample (one output file: the
    reference opens one per leaf-identifier distance, a single one without composite objects).

    ``write(positions, system_length)`` takes the (n, dim) positions of the extracted global state
    in identifier order."""

    def __init__(self, filename):
        super().__init__(filename)
        if len(filename.split(".")) != 2:
            raise AssertionError("the filename must contain one file format suffix")
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

    def write(self, positions, system_length):
        super().write(positions, system_length)
        norms = self.separation_norms(positions, system_length)
        # print(float) writes repr(float): the shortest string that reads back as the same double
        self._files[0].write("".join(repr(float(v)) + "\n" for v in norms))

    def post_run(self):
        for f in self._files:
            f.close()


_HANDLERS = {"separation_output_handler": SeparationOutputHandler, "dummy_output_handler": DummyOutputHandler}


def make_output_handler(cls, options):
    """An output handler from its INI class name (snake case) and section options."""
    if cls not in _HANDLERS:
        raise ConfigurationError("output handler %s is not supported" % cls)
    return _HANDLERS[cls](**options)
