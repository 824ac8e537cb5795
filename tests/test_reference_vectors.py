"""The golden vectors of the reference's own potential unit tests (collected by
tests/golden/make_unit_vectors.py) against the oracle, the reference's own C sources and
the device pair kernels.  Tolerances are the tests' own `places`."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import refapi

VECTORS = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                      "reference_unit_vectors.json")))


def test_vector_count():
    assert len(VECTORS) >= 200
    classes = {v["potential"]["cls"] for v in VECTORS}
    assert classes == set(refapi.POT_CODE)


@pytest.mark.parametrize("math_mode", [0, 1])
def test_oracle_reproduces_reference_unit_vectors(oracle, math_mode):
    bad = []
    for v in VECTORS:
        got = refapi.oracle_eval(oracle, v, math_mode)
        if not refapi.matches(got, v):
            bad.append((v["file"], v["line"], got, v["expected"]))
    assert not bad, bad[:5]


def _ref_or_skip(name):
    from oracle import oracle as O
    lib = O.ref_lib(name)
    if lib is None:
        pytest.skip("reference C sources not compiled (no reference tree)")
    return lib


def test_ipcb_restatement_matches_reference_c(oracle):
    ref = _ref_or_skip("ipcb")
    d = ctypes.c_double
    ref.displacement.restype = d
    ref.displacement.argtypes = [d] * 6
    ref.derivative.restype = d
    ref.derivative.argtypes = [d] * 4
    rng = np.random.default_rng(1)
    L = oracle.lib()
    for _ in range(20000):
        pp = rng.choice([-1, 1]) * rng.uniform(0.1, 3.0)
        s = rng.uniform(-0.5, 0.5, 3)
        dU = rng.exponential(2.0)
        a = ref.displacement(pp, s[0], s[1], s[2], dU, 1.0)
        b = L.ora_ipcb_displacement(pp, s[0], s[1], s[2], dU, 1.0)
        assert a == b or (math.isnan(a) and math.isnan(b))
        assert ref.derivative(pp, *s) == L.ora_ipcb_derivative(pp, *s)


def test_ewald_restatement_matches_reference_c(oracle):
    ref = _ref_or_skip("mic")
    d = ctypes.c_double
    ref.construct_merged_image_coulomb_potential.restype = ctypes.c_void_p
    ref.construct_merged_image_coulomb_potential.argtypes = [ctypes.c_int, ctypes.c_int, d, d]
    ref.derivative.restype = d
    ref.derivative.argtypes = [ctypes.c_void_p, d, d, d]
    ref.destroy_merged_image_coulomb_potential.argtypes = [ctypes.c_void_p]
    rng = np.random.default_rng(2)
    L = oracle.lib()
    for (fc, pc, alpha, length) in [(6, 2, 3.45, 1.0), (6, 2, 3.45, 8.0), (9, 2, 5.0, 2.0)]:
        h_ref = ref.construct_merged_image_coulomb_potential(fc, pc, alpha, length)
        h = L.ora_ewald_create(fc, pc, alpha, length)
        for _ in range(2000):
            s = rng.uniform(-0.5, 0.5, 3) * length
            assert ref.derivative(h_ref, *s) == L.ora_ewald_derivative(h, *s)
        ref.destroy_merged_image_coulomb_potential(h_ref)
        L.ora_ewald_destroy(h)


@pytest.mark.gpu
def test_device_pair_kernels_reproduce_reference_unit_vectors(device, oracle):
    vecs = [v for v in VECTORS if refapi.axis_velocity(v["args"][0]) is not None]
    assert len(vecs) >= 200
    got = refapi.device_eval(vecs)
    bad = [(v["file"], v["line"], g, v["expected"]) for g, v in zip(got, vecs) if not refapi.matches(g, v)]
    assert not bad, bad[:5]
    # and bit for bit the oracle in correctly rounded mode, apart from the Ewald sum whose
    # exp/erfc/cos/sin are the device library's (agreement to 1e-13 relative)
    for g, v in zip(got, vecs):
        want = refapi.oracle_eval(oracle, v, 1)
        if v["potential"]["cls"] == "MergedImageCoulombPotential":
            assert abs(g - want) <= 1e-13 * max(1.0, abs(want))
        else:
            assert g == want or (math.isnan(g) and math.isnan(want)), (v["file"], v["line"], g, want)
