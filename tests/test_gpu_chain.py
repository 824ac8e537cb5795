"""Parity of the plain GPU chain and of the device math against the oracle and the
reference's own event logs.  Integer and event data must agree bit for bit; see
test_oracle_golden.py for how the oracle itself is pinned."""
import os
import random

import numpy as np
import pytest

import systems as S

pytestmark = pytest.mark.gpu

FIELDS = ["tagger", "flags", "active", "target", "tq", "tr", "pos", "ncand", "chash"]


def first_difference(a, b, fields=FIELDS):
    n = min(len(a), len(b))
    first = None
    for f in fields:
        eq = a[f][:n] == b[f][:n]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.flatnonzero(~eq)
        if bad.size and (first is None or bad[0] < first[0]):
            first = (int(bad[0]), f)
    if first is None and len(a) != len(b):
        first = (n, "length")
    return first


def test_mt_words_match_cpython(device):
    import torch
    for seed in (0, 5, 123456789, 2 ** 40 + 3):
        n = 624 * 7
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        device.check(device.lib().ecmc_mt_words(seed, n, device.dptr(out), None))
        words = out.cpu().numpy().view(np.uint32)
        r = random.Random(seed)
        ref = np.array([r.getrandbits(32) for _ in range(n)], dtype=np.uint32)
        assert np.array_equal(words, ref), seed


def test_device_log_pow_correctly_rounded(device, oracle):
    import torch
    rng = np.random.default_rng(3)
    x = np.concatenate([1.0 - rng.random(20000), np.exp(rng.uniform(-40, 40, 20000)), [1.0, 0.5, 2.0, 1e-300]])
    xt = torch.tensor(x, device="cuda")
    out = torch.empty_like(xt)
    device.check(device.lib().ecmc_math_log(len(x), device.dptr(xt), device.dptr(out), None))
    got = out.cpu().numpy()
    want = np.array([oracle.lib().ora_log(v, oracle.MATH_CR) for v in x])
    assert np.array_equal(got, want)
    base = np.concatenate([rng.uniform(0.3, 40.0, 20000), rng.uniform(1e-4, 6.0, 20000)])
    ys = np.concatenate([rng.choice([6.0, 3.0, 14.0, 8.0, 4.5, 2.0, 0.5], 20000),
                         rng.choice([2.0 / 12.0, 2.0 / 6.0, 2.0 / 9.0, 1.0 / 6.0, 1.5], 20000)])
    bt, yt = torch.tensor(base, device="cuda"), torch.tensor(ys, device="cuda")
    out = torch.empty_like(bt)
    device.check(device.lib().ecmc_math_pow(len(base), device.dptr(bt), device.dptr(yt), device.dptr(out), None))
    got = out.cpu().numpy()
    want = np.array([oracle.lib().ora_pow(b, y, oracle.MATH_CR) for b, y in zip(base, ys)])
    assert np.array_equal(got, want)


def _oracle_cr(oracle, fn, *arrays):
    import ctypes
    pd = ctypes.POINTER(ctypes.c_double)
    arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    out = np.empty_like(arrays[0])
    getattr(oracle.lib(), fn)(len(out), *[a.ctypes.data_as(pd) for a in arrays], out.ctypes.data_as(pd),
                              oracle.MATH_CR)
    return out


def test_device_log_pow_correctly_rounded_at_scale(device, oracle):
    # the fast paths (double-length log with a rounding test; k-th root plus one Newton step
    # for 2/p = 1/3, 1/6; the Newton step on t^den = x^num for 2/p = 2/9, 2/7, 1/5; half-integer
    # powers r2^(p/2) for odd p) against quad precision over millions of the arguments the
    # inversion meets: 1 - u for the event draws, (k / (U + dU)) ** (2/p) and r2 ** (p/2)
    import torch
    rng = np.random.default_rng(11)
    x = np.concatenate([1.0 - rng.random(1_000_000), np.exp(rng.uniform(-40, 40, 500_000)),
                        1.0 + rng.uniform(-2 ** -7, 2 ** -7, 500_000)])
    xt = torch.tensor(x, device="cuda")
    out = torch.empty_like(xt)
    print("device log at scale", len(x), flush=True)
    device.check(device.lib().ecmc_math_log(len(x), device.dptr(xt), device.dptr(out), None))
    got = out.cpu().numpy()
    want = _oracle_cr(oracle, "ora_log_n", x)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, "log differs at x = %r" % x[bad[:5]]
    base = np.exp(rng.uniform(-8.0, 22.0, 3_000_000))
    ys = rng.choice([2.0 / 12.0, 2.0 / 6.0, 2.0 / 9.0, 2.0 / 7.0, 2.0 / 10.0, 4.5, 3.5, 0.5 + 7], len(base))
    half = ys > 1.0
    base[half] = np.exp(rng.uniform(-10.0, 10.0, half.sum()))     # r2 ** (p/2) within 2^+-30
    bt, yt = torch.tensor(base, device="cuda"), torch.tensor(ys, device="cuda")
    out = torch.empty_like(bt)
    device.check(device.lib().ecmc_math_pow(len(base), device.dptr(bt), device.dptr(yt), device.dptr(out), None))
    got = out.cpu().numpy()
    want = _oracle_cr(oracle, "ora_pow_n", base, ys)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, "pow differs at x = %r" % base[bad[:5]]


@pytest.mark.parametrize("name", ["hard_disks_n64", "lj_two_factor_n256", "lj_mexican_hat_n108"])
def test_gpu_chain_matches_oracle_on_reference_fixture(device, oracle, name):
    from jellyfysh_amd.chain import DeviceChain
    meta, g = S.load_fixture(name)
    n = len(g["tagger"])
    print("fixture", name, "legs", n, flush=True)
    chain = DeviceChain(S.device_desc(meta, scheduler="list"), g["positions"])
    log = np.concatenate([chain.run(n // 2), chain.run(n - n // 2)])
    # the oracle in correctly rounded mode is what the device's double-double math rounds to
    e = oracle.Engine(S.oracle_config(meta, math_mode=oracle.MATH_CR, scheduler="list"), g["positions"])
    ref = e.run(n)
    assert first_difference(log, ref) is None, first_difference(log, ref)
    assert np.array_equal(chain.positions(), e.positions())
    assert chain.rng_words() == e.rng_words() or meta["factors"][0]["kind"] == "hard_sphere"
    # against the reference's own log: same targets everywhere, times within 1e-12
    assert np.array_equal(log["target"], g["target"])
    assert np.array_equal(log["tagger"], g["tagger"])
    t, tg = log["tq"] + log["tr"], g["tq"] + g["tr"]
    assert np.max(np.abs(t - tg) / np.maximum(np.abs(tg), 1.0)) < 1e-12
    if meta["factors"][0]["kind"] == "hard_sphere":
        assert first_difference(log, g_as_events(g), ["tagger", "target", "tq", "tr", "pos"]) is None


def g_as_events(g):
    out = np.zeros(len(g["tagger"]), dtype=[("tagger", np.int32), ("target", np.int32), ("tq", np.float64),
                                            ("tr", np.float64), ("pos", np.float64, (3,))])
    for f in out.dtype.names:
        out[f] = g[f]
    return out


def test_gpu_hard_disks_config1_bit_exact(device, oracle):
    # configuration 1: N=256, phi=0.60, list scheduler, chains of 0.5
    from jellyfysh_amd.chain import DeviceChain
    meta, pos = S.hard_disk_system(n=256, phi=0.60, seed=7)
    legs = 200000
    print("hard disks N=256 legs", legs, flush=True)
    chain = DeviceChain(S.device_desc(meta), pos)
    log = chain.run(legs)
    e = oracle.Engine(S.oracle_config(meta), pos)
    ref = e.run(legs)
    assert first_difference(log, ref) is None, first_difference(log, ref)
    assert np.array_equal(chain.positions(), e.positions())
    assert np.array_equal(chain.estimators()[:6], e.estimators()[:6])


def test_gpu_lj_cells_bit_exact_against_cr_oracle(device, oracle):
    # LJ 12-6 as two inverse-power factors, 4^3 FCC cells (N=256), reference wiring
    from jellyfysh_amd.chain import DeviceChain
    meta, pos = S.lj_system(n_cells=4, rho=0.8, seed=11, cells_per_side=4)
    legs = 6000
    print("LJ N=256 legs", legs, flush=True)
    chain = DeviceChain(S.device_desc(meta), pos)
    log = chain.run(legs)
    e = oracle.Engine(S.oracle_config(meta, math_mode=oracle.MATH_CR), pos)
    ref = e.run(legs)
    assert first_difference(log, ref) is None, first_difference(log, ref)
    assert chain.rng_words() == e.rng_words()


def test_gpu_lj_truncated_bit_exact_against_cr_oracle(device, oracle):
    from jellyfysh_amd.chain import DeviceChain
    meta, pos = S.lj_system(n_cells=5, rho=0.8, seed=4, cutoff=2.5)
    legs = 2500
    print("LJ truncated N=500 legs", legs, flush=True)
    chain = DeviceChain(S.device_desc(meta), pos)
    log = chain.run(legs)
    e = oracle.Engine(S.oracle_config(meta, math_mode=oracle.MATH_CR), pos)
    ref = e.run(legs)
    assert first_difference(log, ref) is None, first_difference(log, ref)


@pytest.mark.gpu
def test_gpu_leg_loop_kernel_variants_cover_the_systems(device):
    """Each system runs the first specialized leg-loop kernel whose features cover its taggers'
    (ecmc_system_kernel_variant): LJ as truncated inverse powers the inverse-power variant, the
    charged soft spheres the Coulomb variant, the flexible dipoles the factor-map variant, 2D hard
    disks the full kernel; the parity tests of each workload run through these kernels."""
    from jellyfysh_amd.chain import DeviceChain
    from jellyfysh_amd.config import RunConfig
    from jellyfysh_amd.workloads import charged_soft_spheres, make_chain
    IP, IPT, DEP, HS = device.POT_INVERSE_POWER, device.POT_INVERSE_POWER_TRUNC, device.POT_DISPLACED_EVEN_POWER, \
        device.POT_HARD_SPHERE
    COUL, FMAP, CB, ALL = 1 << 9, 1 << 10, 1 << 11, (1 << 13) - 1

    def variant(chain):
        try:
            return chain.kernel_variant()
        finally:
            chain.close()

    meta, pos = S.lj_system(n_cells=3, cutoff=2.5)
    assert variant(DeviceChain(S.device_desc(meta), pos)) == (1 << IPT, (1 << IP) | (1 << IPT))
    f, v = variant(make_chain(charged_soft_spheres(n=32, system_length=4.0, seed=1, cells_per_side=4)))
    assert f == (1 << IP) | COUL | CB and v == f
    cfg = RunConfig.from_file(os.path.join(os.path.dirname(S.__file__), "..", "jellyfysh_amd", "configs",
                                           "flexible_dipoles.ini"))
    p, q, w = cfg.initial_state(0)
    f, v = variant(DeviceChain(cfg.desc(len(p), 0, rng_skip=w), p, q))
    assert f == (1 << IP) | (1 << DEP) | COUL | FMAP and v == f
    meta, pos = S.hard_disk_system(n=64)
    assert variant(DeviceChain(S.device_desc(meta), pos)) == (1 << HS, ALL)
