"""Output handlers (jellyfysh_amd/output.py): the reference's text writer with the run hash, the
separation output handler fed from the device chain's sampling events, and their INI wiring.

Pins: the separation norms of the device kernel equal the oracle's restatement of
SeparationOutputHandler.write (separation_output_handler.py:81-95, hypercubic_setting.py:120-172,
vectors.py:29-43) bit for bit; a sampled run of the hard-disk INI writes, sample after sample,
the norms of the oracle's positions at the same sampling events."""
import os

import numpy as np
import pytest

import systems as S

CONFIGS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jellyfysh_amd", "configs")


def sampled_hard_disk_ini(filename, end_time=2.0, interval=0.37):
    text = open(os.path.join(CONFIGS, "hard_disks.ini")).read()
    text = text.replace("    start_of_run (no_in_state_tagger)",
                        "    start_of_run (no_in_state_tagger),\n    sampling (no_in_state_tagger)")
    text = text.replace("end_of_run_time = 1.0e9", "end_of_run_time = %r" % end_time)
    text = text.replace("trash = end_of_chain, nearby_sphere, cell_boundary, start_of_run, end_of_run",
                        "trash = end_of_chain, nearby_sphere, cell_boundary, start_of_run, end_of_run, sampling")
    text = text.replace("create = nearby_sphere, cell_boundary, end_of_chain, end_of_run\ntrash = start_of_run",
                        "create = nearby_sphere, cell_boundary, end_of_chain, end_of_run, sampling\n"
                        "trash = start_of_run")
    text += ("\n[Sampling]\ncreate = sampling\ntrash = sampling\n"
             "event_handler = fixed_interval_sampling_event_handler\n\n"
             "[FixedIntervalSamplingEventHandler]\nsampling_interval = %r\n"
             "output_handler = separation_output_handler\n\n"
             "[InputOutputHandler]\noutput_handlers = separation_output_handler\n\n"
             "[SeparationOutputHandler]\nfilename = %s\n" % (interval, filename))
    return text


def test_text_writer_header_and_rename(tmp_path):
    from jellyfysh_amd.output import HardBufferedTextWriter, get_uuid
    f = str(tmp_path / "Samples.dat")
    w = HardBufferedTextWriter(f)
    print(0.25, file=w)
    assert os.path.exists(f + ".tmp") and not os.path.exists(f)
    w.close()
    assert not os.path.exists(f + ".tmp")
    lines = open(f).read().splitlines()
    assert lines == ["# Run identification hash: %s" % get_uuid(), "0.25"]
    assert get_uuid() == get_uuid()


def test_output_handler_checks_its_directory(tmp_path):
    from jellyfysh_amd.output import ConfigurationError, DummyOutputHandler, SeparationOutputHandler, \
        make_output_handler
    with pytest.raises(ConfigurationError):
        SeparationOutputHandler(str(tmp_path / "missing" / "Samples.dat"))
    with pytest.raises(ConfigurationError):
        make_output_handler("dcd_output_handler", {"filename": "x.dcd"})
    d = make_output_handler("dummy_output_handler", {})
    assert isinstance(d, DummyOutputHandler)
    d.write(np.zeros((2, 2)), 1.0)
    d.post_run()


def test_ini_wires_the_sampling_handler_to_its_output_handler(tmp_path):
    from jellyfysh_amd.config import RunConfig
    cfg = RunConfig.from_string(sampled_hard_disk_ini(str(tmp_path / "Separations.dat")))
    assert cfg.output_handlers == {"separation_output_handler":
                                   ("separation_output_handler", {"filename": str(tmp_path / "Separations.dat")})}
    (s,) = [t for t in cfg.taggers if t["kind"] == "sampling"]
    assert s["output_handler"] == "separation_output_handler" and s["sampling_interval"] == 0.37


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_device_separation_norms_equal_the_oracle(device, oracle, dim):
    from jellyfysh_amd.output import SeparationOutputHandler
    rng = np.random.default_rng(dim)
    L = 3.7
    pos = rng.uniform(0.0, L, (300, dim))
    # separations of exactly +-L/2 and of nearly L/2 along an axis, where the wrap decides
    pos[1] = pos[0]
    pos[1, 0] = (pos[0, 0] + L / 2) % L
    pos[2] = pos[0]
    pos[2, 0] = np.nextafter((pos[0, 0] + L / 2) % L, 0.0)
    print("separation norms, dim", dim, flush=True)
    got = SeparationOutputHandler.separation_norms(pos, L)
    want = np.empty_like(got)
    import ctypes
    pd = ctypes.POINTER(ctypes.c_double)
    p = np.ascontiguousarray(pos)
    oracle.lib().ora_separation_norms(dim, len(p), p.ctypes.data_as(pd), L, want.ctypes.data_as(pd), oracle.MATH_CR)
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_sampled_run_writes_the_oracle_separations(device, oracle, tmp_path):
    import ctypes
    from jellyfysh_amd.config import RunConfig, SingleProcessMediator
    fname = str(tmp_path / "Separations.dat")
    cfg = RunConfig.from_string(sampled_hard_disk_ini(fname))
    meta, pos = S.hard_disk_system(n=256, seed=7)
    print("sampled hard-disk run to t = 2", flush=True)
    med = SingleProcessMediator(cfg, pos, seed=7)
    legs = med.run()
    lines = open(fname).read().splitlines()
    assert lines[0].startswith("# Run identification hash: ")
    got = np.array([float(v) for v in lines[1:]])
    # the oracle, leg by leg, with the same wiring: the norms at each of its sampling events
    ref = oracle.Engine(S.oracle_config_from_wiring(cfg.dim, 256, cfg.system_length, cfg.beta, cfg.cells_per_side,
                                                    cfg.taggers, 7, cfg.scheduler, oracle.MATH_CR,
                                                    cfg.max_occupants), pos)
    sampling = [i for i, t in enumerate(cfg.taggers) if t["kind"] == "sampling"][0]
    pd = ctypes.POINTER(ctypes.c_double)
    want, n_legs = [], 0
    while True:
        ev = ref.run(1)
        if len(ev) == 0:
            break
        n_legs += 1
        if int(ev["tagger"][0]) == sampling:
            p = np.ascontiguousarray(ref.positions())
            out = np.empty(256 * 255 // 2)
            oracle.lib().ora_separation_norms(2, 256, p.ctypes.data_as(pd), cfg.system_length,
                                              out.ctypes.data_as(pd), oracle.MATH_CR)
            want.append(out)
    assert n_legs == legs
    assert len(want) == 5                      # samples at 0.37, 0.74, ..., 1.85
    assert np.array_equal(got, np.concatenate(want))
