"""Device parity at the benchmark configurations' own sizes.

- Plain LJ 12-6 at N = 4096, rho = 0.8, 6^3 cells, heap scheduler (the plain workload of the
  benchmark): the device event log against the oracle in correctly rounded mode, bit for bit,
  for the truncated-and-shifted factors (rc = 2.5) and for the reference's wiring without
  cutoff (the two InversePowerPotential factors over the 27 nearby cells).
- Charged soft spheres at N = 512, L = 8, 8^3 cells: the far cells through the cell-bounding
  potential and through the cell veto, against the oracle (targets, taggers and candidate
  counts equal; event times within 1e-12, the device's erfc/exp in the Ewald tables rounding
  differently from glibc's).
- The batched workload of the benchmark, all 4096 replicas, one step of 1000 legs under the
  bound audit (every candidate of the 27 nearby cells, in culled cells too, whose exact
  displacement is within the horizon must be a contender), the same step in the default
  filtered mode ending in the same state, and 16 sampled replicas against the batched oracle.

The oracle is pinned separately (reference event logs, reference unit-test vectors, the
reference's C files, CPython's random streams; see test_oracle_golden.py and
test_reference_vectors.py)."""
import numpy as np
import pytest

import systems as S

pytestmark = pytest.mark.gpu

FIELDS = ["tagger", "flags", "active", "target", "tq", "tr", "pos", "ncand", "chash"]


def _first_difference(a, b, fields=FIELDS):
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


def _lj_n4096(cutoff):
    from jellyfysh_amd.batch import lattice_positions
    n = 4096
    L = (n / 0.8) ** (1 / 3)
    pos = lattice_positions(n, L, 0.01, np.random.default_rng(7))
    if cutoff:
        factors = [dict(tag="lj_twelve", kind="inverse_power_truncated", power=12.0, prefactor=4.0, cutoff=cutoff),
                   dict(tag="lj_six", kind="inverse_power_truncated", power=6.0, prefactor=-4.0, cutoff=cutoff)]
    else:
        factors = [dict(tag="lj_twelve", kind="inverse_power", power=12.0, prefactor=4.0),
                   dict(tag="lj_six", kind="inverse_power", power=6.0, prefactor=-4.0)]
    meta = dict(dim=3, n=n, system_length=L, beta=1 / 1.5, cells_per_side=6, chain_time=1.0, end_time=1e9,
                scheduler="heap", seed=5, sampling_interval=0.0, factors=factors)
    return meta, pos


@pytest.mark.parametrize("cutoff", [2.5, 0.0])
def test_gpu_plain_lj_n4096_bit_exact(device, oracle, cutoff):
    from jellyfysh_amd.chain import DeviceChain
    meta, pos = _lj_n4096(cutoff)
    legs = 3000
    print("plain LJ N = 4096, 6^3 cells, cutoff %g, %d legs" % (cutoff, legs), flush=True)
    chain = DeviceChain(S.device_desc(meta), pos)
    log = np.concatenate([chain.run(legs // 2), chain.run(legs - legs // 2)])
    e = oracle.Engine(S.oracle_config(meta, math_mode=oracle.MATH_CR), pos)
    ref = e.run(legs)
    assert _first_difference(log, ref) is None, _first_difference(log, ref)
    assert chain.rng_words() == e.rng_words()
    assert np.array_equal(chain.positions(), e.positions())
    assert np.array_equal(chain.estimators()[:6], e.estimators()[:6])
    assert log["ncand"].mean() > 900          # about 2 x 511 pair candidates per leg


@pytest.mark.parametrize("far", ["cell_bounding", "cell_veto"])
def test_gpu_charged_soft_spheres_n512_match_oracle(device, oracle, far):
    from jellyfysh_amd.workloads import charged_soft_spheres, make_chain
    sysd = charged_soft_spheres(seed=3, far=far)
    legs = 1500
    print("charged soft spheres N = 512, far cells through %s, %d legs" % (far, legs), flush=True)
    chain = make_chain(sysd)
    got = chain.run(legs)
    cfg = S.oracle_config_from_wiring(3, 512, sysd["system_length"], sysd["beta"], sysd["cells_per_side"],
                                      sysd["taggers"], sysd["seed"], scheduler="heap", math_mode=oracle.MATH_CR,
                                      max_occupants=1)
    e = oracle.Engine(cfg, sysd["positions"], sysd["charges"])
    want = e.run(legs)
    assert len(got) == len(want) == legs
    for f in ("tagger", "flags", "active", "target", "ncand"):
        bad = np.flatnonzero(got[f] != want[f])
        assert bad.size == 0, "field %s differs first at leg %d" % (f, bad[0])
    t_got, t_want = got["tq"] + got["tr"], want["tq"] + want["tr"]
    assert np.max(np.abs(t_got - t_want) / np.maximum(np.abs(t_want), 1.0)) < 1e-12
    assert chain.rng_words() == e.rng_words()
    dpos = np.abs(chain.positions() - e.positions())
    assert np.max(np.minimum(dpos, 8.0 - dpos)) < 1e-9
    counts = np.bincount(got["tagger"], minlength=len(sysd["taggers"]))
    assert counts[0] > 0                      # the far-cell handler commits events


def test_gpu_batched_full_workload_under_audit(device, oracle):
    """All 4096 replicas of the benchmark's workload, one step of 1000 legs in its default size
    classes: no bound violation (culled cells included), the filtered mode ends in the same
    state, and 16 replicas (the smallest and the largest N, the densest N = 2048, the rest spread
    over N) equal the batched oracle event for event."""
    from jellyfysh_amd.batch import DeviceBatch, batched_lj_workload, default_classes
    n, lengths, pos = batched_lj_workload(4096, seed=2024)
    legs = 1000
    print("batched workload, 4096 replicas, %d legs, audit mode" % legs, flush=True)
    audited = DeviceBatch(n, lengths, pos, seed=17, classes=default_classes())
    audited.set_mode(2)
    audited.set_log(legs)
    audited.run(legs)
    assert audited.audit() == 0
    logs = audited.log()
    sa = audited.states()
    pa = audited.positions()
    audited.close()
    print("batched workload, filtered mode", flush=True)
    plain = DeviceBatch(n, lengths, pos, seed=17, classes=default_classes())
    plain.run(legs)
    sp = plain.states()
    pp = plain.positions()
    plain.close()
    for k in ("active", "direction", "legs", "chain_left", "est"):
        assert np.array_equal(sa[k], sp[k]), k
    assert all(np.array_equal(a, b) for a, b in zip(pa, pp))
    assert np.all(sa["legs"] == legs)
    rho = n / lengths ** 3
    big = np.flatnonzero(n == n.max())
    picks = {int(np.argmin(n)), int(np.argmax(n)), int(big[np.argmax(rho[big])]), int(np.argmax(rho)),
             int(np.argmin(rho))}
    order = np.argsort(n, kind="stable")
    for q in np.linspace(0, len(n) - 1, 16).astype(int):
        if len(picks) >= 16:
            break
        picks.add(int(order[q]))
    print("oracle replicas", sorted(picks), flush=True)
    for r in sorted(picks):
        rep = oracle.BatchedReplica(pos[r], lengths[r], replica=r, seed=17)
        want = rep.run(legs)
        for f in ("kind", "active", "target", "factor", "displacement", "x"):
            bad = np.flatnonzero(logs[r][f] != want[f])
            assert bad.size == 0, (r, int(n[r]), f, int(bad[0]))
        assert np.array_equal(pa[r], rep.pos)
        assert np.array_equal(sa["est"][r][:6], np.array(rep.state.est)[:6])
