"""ecmc_systems_run_noalloc: the launch of independent systems without any allocation (its
arguments in the library's fixed device array) ends every system exactly where
ecmc_systems_run ends the same systems, on two calls in a row, the second with fewer systems,
and rejects more systems than its array holds."""
import ctypes

import numpy as np
import pytest

import systems as S

pytestmark = pytest.mark.gpu


def _state(c):
    return c.positions(), c.estimators(), c.rng_words(), c.active()


def _same(a, b):
    for x, y in zip(_state(a), _state(b)):
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y)
        else:
            assert x == y


def _chains(seeds):
    from jellyfysh_amd.chain import DeviceChain, make_desc
    from jellyfysh_amd.workloads import charged_soft_spheres, make_chain
    out = []
    for s in seeds:
        if s % 2:
            out.append(make_chain(charged_soft_spheres(n=32, system_length=4.0, seed=s, cells_per_side=4)))
        else:
            meta, pos = S.lj_system(n_cells=3, rho=0.8, seed=s)
            out.append(DeviceChain(S.device_desc(meta), pos))
    return out


def test_gpu_noalloc_launch_equals_the_allocating_launch(device):
    from jellyfysh_amd import _lib
    from jellyfysh_amd.chain import run_together
    print("no-allocation launch: 6 systems, then 3", flush=True)
    seeds = [1, 2, 3, 4, 5, 6]
    a, b = _chains(seeds), _chains(seeds)
    na, _ = run_together(a, 700)
    nb, _ = run_together(b, 700, noalloc=True)
    assert np.array_equal(na, nb) and np.all(na == 700)
    for x, y in zip(a, b):
        _same(x, y)
    # second call, fewer systems (the array keeps the first call's arguments beyond them)
    na, _ = run_together(a[3:], 400)
    nb, _ = run_together(b[3:], 400, noalloc=True)
    assert np.array_equal(na, nb) and np.all(na == 400)
    for x, y in zip(a, b):
        _same(x, y)
    # more systems than the array holds: rejected before anything runs
    hs = (ctypes.c_void_p * (_lib.SYSTEMS_NOALLOC_MAX + 1))(*([b[0]._h.value] * (_lib.SYSTEMS_NOALLOC_MAX + 1)))
    n = np.zeros(_lib.SYSTEMS_NOALLOC_MAX + 1, dtype=np.int64)
    rc = _lib.lib().ecmc_systems_run_noalloc(hs, _lib.SYSTEMS_NOALLOC_MAX + 1, 10,
                                             n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    assert rc == -1 and b"ECMC_SYSTEMS_NOALLOC_MAX" in _lib.lib().ecmc_last_error()
    with pytest.raises(Exception):
        run_together([b[0], b[0]], 10, noalloc=True)
    for c in a + b:
        c.close()
