"""The cell bounding potential (cell_bounding_potential.py, inner_point_estimator.py) with
its tagger and event handler, the surplus tagger and the bounded cell occupancy.

Pins: the estimator's tables bound the merged-image Coulomb derivative over every cell
separation (random points of each region); one table entry recomputed from the reference's
own C derivative on the estimator's grid; the chain runs the reference's cell-bounded wiring
with every tagger firing; the device commits the oracle's events bit for bit."""
import math
import os

import numpy as np
import pytest

import systems as S


def _table(oracle, cps=(3, 5, 7), L=1.0):
    cells = oracle.Cells(3, L, list(cps))
    return cells, cells.cell_bound_table(L, 3.45, 6, 2)


def test_cell_bound_table_bounds_the_derivative(oracle):
    cells, tb = _table(oracle)
    lib = oracle.lib()
    h = lib.ora_ewald_create(6, 2, 3.45, 1.0)
    rng = np.random.default_rng(0)
    zmin, zmax = cells.bounds(0)
    far = 0
    try:
        for c in range(cells.count()):
            if np.isnan(tb[c, 0, 0]):
                assert c in cells.nearby(0)
                continue
            far += 1
            mn, mx = cells.bounds(c)
            lo, hi = mn - zmax, mx - zmin
            for _ in range(40):
                s = rng.uniform(lo, hi)
                s = (s + 0.5) % 1.0 - 0.5
                for d in range(3):
                    p = [s[(d + k) % 3] for k in range(3)]
                    der = lib.ora_ewald_derivative(h, *p)
                    assert tb[c, d, 1] <= der <= tb[c, d, 0], (c, d, der, tb[c, d])
    finally:
        lib.ora_ewald_destroy(h)
    assert far == cells.count() - 27


def test_cell_bound_table_entry_from_the_reference_c(oracle):
    ref = oracle.ref_lib("mic")
    if ref is None:
        pytest.skip("the reference's merged-image Coulomb C file is not built here")
    import ctypes
    cells, tb = _table(oracle)
    ref.construct_merged_image_coulomb_potential.restype = ctypes.c_void_p
    ref.construct_merged_image_coulomb_potential.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                             ctypes.c_double]
    ref.derivative.restype = ctypes.c_double
    ref.derivative.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    h = ref.construct_merged_image_coulomb_potential(6, 2, 3.45, 1.0)
    zmin, zmax = cells.bounds(0)
    for c, d in [(52, 0), (77, 2), (100, 1)]:
        if np.isnan(tb[c, 0, 0]):
            continue
        mn, mx = cells.bounds(c)
        lo, hi = list(mn - zmax), list(mx - zmin)
        up, low = -math.inf, math.inf
        for ix in range(11):
            px = lo[0] + (hi[0] - lo[0]) * ix / 10
            for iy in range(11):
                py = lo[1] + (hi[1] - lo[1]) * iy / 10
                for iz in range(11):
                    pz = lo[2] + (hi[2] - lo[2]) * iz / 10
                    sep = [(v + 0.5) % 1.0 - 0.5 for v in (px, py, pz)]
                    p = [sep[(d + k) % 3] for k in range(3)]
                    der = 1.0 * 1.0 * 1.0 * ref.derivative(h, *p)
                    up, low = max(up, der), min(low, der)
        up = up * 1.5 if up > 0 else up / 1.5
        low = low / 1.5 if low > 0 else low * 1.5
        assert (up, low) == (tb[c, d, 0], tb[c, d, 1])


def _oracle_engine(oracle, meta, taggers, pos, charges, math_mode):
    cfg = S.oracle_config_from_wiring(3, meta["n"], meta["system_length"], meta["beta"], meta["cells_per_side"],
                                      taggers, meta["seed"], scheduler="heap", math_mode=math_mode,
                                      max_occupants=1)
    return oracle.Engine(cfg, pos, charges)


def test_oracle_cell_bounded_run(oracle):
    meta, taggers, pos, charges = S.coulomb_cell_bounded_system(n=12)
    e = _oracle_engine(oracle, meta, taggers, pos, charges, oracle.MATH_LIBM)
    log = e.run(20000)
    assert len(log) == 20000
    counts = np.bincount(log["tagger"], minlength=len(taggers))
    for i, t in enumerate(taggers[:5]):
        assert counts[i] > 0, t["tag"]
    lifts = np.count_nonzero(log["flags"] & 1)
    est = e.estimators()
    assert est[6] > 0 and lifts > 0      # thinning rejects and accepts


@pytest.mark.gpu
@pytest.mark.parametrize("n", [2, 12])
def test_gpu_cell_bounded_matches_oracle(device, oracle, n):
    from jellyfysh_amd.chain import DeviceChain, make_desc
    meta, taggers, pos, charges = S.coulomb_cell_bounded_system(n=n)
    legs = 6000
    print("cell-bounded chain, n = %d, %d legs" % (n, legs), flush=True)
    desc = make_desc(3, n, meta["system_length"], meta["beta"], meta["cells_per_side"], taggers, meta["seed"],
                     scheduler="heap", max_occupants=1)
    got = DeviceChain(desc, pos, charges).run(legs)
    want = _oracle_engine(oracle, meta, taggers, pos, charges, oracle.MATH_CR).run(legs)
    assert len(got) == len(want) == legs
    # The event sequence is exact.  The bound tables come from erfc/exp of the device's math
    # library on one side and of glibc on the other (neither correctly rounded), so event times
    # and positions agree to rounding, within 1e-12 relative.
    for f in ("tagger", "flags", "target", "ncand", "chash"):
        bad = np.flatnonzero(got[f] != want[f])
        assert bad.size == 0, "field %s differs first at leg %d" % (f, bad[0])
    t_got, t_want = got["tq"] + got["tr"], want["tq"] + want["tr"]
    rel_t = np.max(np.abs(t_got - t_want) / np.maximum(np.abs(t_want), 1.0))
    rel_p = np.max(np.abs(got["pos"] - want["pos"])) / meta["system_length"]
    print("largest relative deviation: time %.2e, position %.2e" % (rel_t, rel_p), flush=True)
    assert rel_t < 1e-12 and rel_p < 1e-12
