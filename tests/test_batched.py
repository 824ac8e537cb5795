"""The batched-replica form: the oracle's restatement on the CPU (Philox KATs, invariants,
physics sanity) and the device kernel against it (bit-exact event logs with the FP32
filter, the all-FP64 mode and the bound audit)."""
import numpy as np
import pytest

# Random123 known-answer vectors of Philox4x32-10
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def test_philox_known_answers(oracle):
    for ctr, key, want in PHILOX_KAT:
        assert oracle.philox(ctr, key) == want


def small_batch(seed=5):
    from jellyfysh_amd.batch import lattice_positions
    rng = np.random.default_rng(seed)
    ns = [64, 110, 300, 700, 1500]
    rhos = [0.9, 0.7, 0.8, 0.85, 0.75]
    lengths = [(n / r) ** (1 / 3) for n, r in zip(ns, rhos)]
    pos = [lattice_positions(n, l, 0.01, rng) for n, l in zip(ns, lengths)]
    return ns, lengths, pos


def test_batched_oracle_invariants(oracle):
    ns, lengths, pos = small_batch()
    for r in (0, 2):
        rep = oracle.BatchedReplica(pos[r], lengths[r], replica=r, seed=9)
        log = rep.run(4000)
        assert np.all(log["displacement"] >= 0)
        assert np.all((rep.pos >= 0) & (rep.pos < lengths[r]))
        lifts = log["kind"] == 0
        assert lifts.sum() > 100
        # a lift hands the activity on to its target
        assert np.array_equal(log["active"][1:][lifts[:-1]], log["target"][:-1][lifts[:-1]])
        # chains: displacement between chain ends sums to the chain length
        ends = np.flatnonzero(log["kind"] == 2)
        for a, b in zip(ends[:-1], ends[1:]):
            assert abs(log["displacement"][a + 1:b + 1].sum() - 1.0) < 1e-12


def virial_pressure(pos, L, beta, rc):
    """beta P / rho = 1 + beta / (3N) sum_{i<j, r<rc} r F(r) for LJ 12-6 (eps = sigma = 1);
    the shift of the potential does not change the forces."""
    n = len(pos)
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)[np.triu_indices(n, 1)]
    r2 = r2[r2 < rc * rc]
    s6 = r2 ** -3
    return 1.0 + beta / (3 * n) * np.sum(48.0 * (s6 * s6 - 0.5 * s6))


@pytest.mark.slow
def test_batched_oracle_pressure_matches_virial(oracle):
    # The ECMC estimator (lifting separations over chain displacement) and the virial of
    # the sampled configurations are two independent routes to the same pressure.
    from jellyfysh_amd.batch import lattice_positions
    rng = np.random.default_rng(1)
    n, beta, rc = 256, 1 / 1.5, 2.5
    L = (n / 0.8) ** (1 / 3)
    rep = oracle.BatchedReplica(lattice_positions(n, L, 0.01, rng), L, replica=0, seed=3, beta=beta, cutoff=rc,
                                math_mode=oracle.MATH_LIBM)
    rep.run(150000, log=False)
    est0 = np.array(rep.state.est)
    vir = []
    for _ in range(120):
        rep.run(5000, log=False)
        vir.append(virial_pressure(rep.pos, L, beta, rc))
    est = np.array(rep.state.est) - est0
    p_ecmc = 1.0 + est[0] / est[1]
    blocks = np.array(vir).reshape(24, 5).mean(axis=1)
    err = blocks.std(ddof=1) / np.sqrt(len(blocks))
    assert 2.0 < p_ecmc < 4.5
    assert abs(p_ecmc - np.mean(vir)) < 4 * err + 0.05, (p_ecmc, np.mean(vir), err)


def _device_batch(classes, mode, log_legs, seed=9):
    from jellyfysh_amd.batch import DeviceBatch
    ns, lengths, pos = small_batch()
    b = DeviceBatch(ns, lengths, pos, seed=seed, classes=classes)
    b.set_mode(mode)
    b.set_log(log_legs)
    return b, ns, lengths, pos


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_batched_matches_oracle(device, oracle, mode):
    legs = 1500
    classes = ([128, 512, 4096], [1, 2, 4])
    print("batched mode", mode, "legs", legs, flush=True)
    b, ns, lengths, pos = _device_batch(classes, mode, legs)
    b.run(legs // 2)
    b.run(legs - legs // 2)
    got = b.log()
    st = b.states()
    for r in range(len(ns)):
        rep = oracle.BatchedReplica(pos[r], lengths[r], replica=r, seed=9)
        want = rep.run(legs)
        for f in ("kind", "active", "target", "factor", "displacement", "x"):
            bad = np.flatnonzero(got[r][f] != want[f])
            assert bad.size == 0, (r, f, int(bad[0]), got[r][bad[0]], want[bad[0]])
        assert np.array_equal(st["est"][r][:6], np.array(rep.state.est)[:6])
        assert st["active"][r] == rep.state.active
    allpos = b.positions()
    for r in range(len(ns)):
        rep = oracle.BatchedReplica(pos[r], lengths[r], replica=r, seed=9)
        rep.run(legs, log=False)
        assert np.array_equal(allpos[r], rep.pos)


@pytest.mark.gpu
@pytest.mark.parametrize("classes", [([128, 512, 4096], [1, 2, 4]), ([16384], [1])])
def test_gpu_batched_bounds_never_exceed_exact(device, classes):
    b, ns, lengths, pos = _device_batch(classes, 2, 0)
    print("batched audit", classes, flush=True)
    b.run(800)
    assert b.audit() == 0
    st = b.states()["stats"]
    # the filter leaves few exact inversions per leg
    assert st[:, 3].sum() < 0.1 * st[:, 0].sum()


@pytest.mark.gpu
def test_gpu_batched_graph_replay_matches_direct_launches(device):
    classes = ([128, 512, 4096], [1, 1, 2])
    b1, *_ = _device_batch(classes, 0, 0, seed=21)
    b2, *_ = _device_batch(classes, 0, 0, seed=21)
    print("batched graph replay", flush=True)
    for _ in range(3):
        b1.run(200)
    b2.run_graph(200, 3)
    s1, s2 = b1.states(), b2.states()
    for k in ("active", "direction", "legs"):
        assert np.array_equal(s1[k], s2[k])
    assert np.array_equal(s1["est"], s2["est"])
    for p1, p2 in zip(b1.positions(), b2.positions()):
        assert np.array_equal(p1, p2)


@pytest.mark.gpu
def test_gpu_batched_class_timing(device):
    # per-class events on the class streams time direct launches without changing the result;
    # graph replays are not class-timed
    classes = ([128, 512, 4096], [1, 1, 2])
    b1, *_ = _device_batch(classes, 0, 0, seed=23)
    b2, *_ = _device_batch(classes, 0, 0, seed=23)
    print("batched class timing", flush=True)
    b1.run_graph(150, 3)
    b2.set_class_timing(True)
    counts = np.bincount(np.searchsorted(np.asarray(b2.classes[0]), b2.n, side="left"), minlength=3)
    for _ in range(3):
        total = b2.run(150)
        for (ms, k), cnt in zip(b2.class_ms(), counts):
            assert k == (1 if cnt else 0)
            assert (ms > 0) == bool(cnt)
            assert ms <= total * 1.01
    assert np.array_equal(b1.states()["est"], b2.states()["est"])
    b2.run_graph(150, 1)
    assert all(k == 0 for _, k in b2.class_ms())


@pytest.mark.gpu
def test_gpu_batched_chunks_migrate_between_blocks(device):
    # two workgroups share five replicas: after every chunk a replica may continue on another
    # workgroup (another XCD), and the result equals one workgroup per replica, bit for bit
    classes = ([4096], [1])
    b1, *_ = _device_batch(classes, 0, 0, seed=29)
    b2, *_ = _device_batch(classes, 0, 0, seed=29)
    b2.set_max_blocks(2)
    print("batched chunk migration", flush=True)
    for b in (b1, b2):
        b.run(700)
        b.run_graph(300, 2)
    s1, s2 = b1.states(), b2.states()
    for k in ("active", "direction", "legs", "stats"):
        assert np.array_equal(s1[k], s2[k])
    assert np.array_equal(s1["est"], s2["est"])
    for p1, p2 in zip(b1.positions(), b2.positions()):
        assert np.array_equal(p1, p2)


@pytest.mark.gpu
@pytest.mark.parametrize("chain_length", [1.0, 4.0])
def test_gpu_batched_one_wavefront_per_replica_matches_oracle(device, oracle, chain_length):
    """The benchmark's layout (one size class, every replica on one wavefront) against the
    oracle, for boxes without cells (N 64, 110, 300), with three cells per side, where the
    nearby cells are the whole box (N 700), and with five (N 1500).  Chains of 4.0 reach half
    the box, which turns the minimum-image ambiguity tests on in the boxes with cells."""
    from jellyfysh_amd.batch import DeviceBatch
    ns, lengths, pos = small_batch()
    legs = 1200
    print("one wavefront per replica, chain length", chain_length, flush=True)
    b = DeviceBatch(ns, lengths, pos, seed=9, classes=([16384], [1]), chain_length=chain_length)
    b.set_log(legs)
    b.run(legs // 3)
    b.run(legs - legs // 3)
    got = b.log()
    st = b.states()
    allpos = b.positions()
    for r in range(len(ns)):
        rep = oracle.BatchedReplica(pos[r], lengths[r], replica=r, seed=9, chain_length=chain_length)
        want = rep.run(legs)
        for f in ("kind", "active", "target", "factor", "displacement", "x"):
            bad = np.flatnonzero(got[r][f] != want[f])
            assert bad.size == 0, (r, f, int(bad[0]), got[r][bad[0]], want[bad[0]])
        assert np.array_equal(st["est"][r][:6], np.array(rep.state.est)[:6])
        assert st["active"][r] == rep.state.active
        assert np.array_equal(allpos[r], rep.pos)
    b.close()


@pytest.mark.gpu
def test_gpu_batched_largest_replicas_match_oracle(device, oracle):
    """The workload's largest replicas (N = 2048 at rho 0.7 and 0.9: 5^3 cells, the most LDS per
    replica) in the benchmark's layout, bit-exact against the oracle."""
    from jellyfysh_amd.batch import DeviceBatch, lattice_positions
    rng = np.random.default_rng(12)
    ns = [2048, 2048]
    lengths = [(2048 / 0.7) ** (1 / 3), (2048 / 0.9) ** (1 / 3)]
    pos = [lattice_positions(n, l, 0.01, rng) for n, l in zip(ns, lengths)]
    legs = 400
    print("two replicas of N = 2048, %d legs" % legs, flush=True)
    b = DeviceBatch(ns, lengths, pos, seed=4, classes=([16384], [1]))
    b.set_log(legs)
    b.run(legs)
    got = b.log()
    allpos = b.positions()
    for r in range(2):
        rep = oracle.BatchedReplica(pos[r], lengths[r], replica=r, seed=4)
        want = rep.run(legs)
        for f in ("kind", "active", "target", "factor", "displacement", "x"):
            bad = np.flatnonzero(got[r][f] != want[f])
            assert bad.size == 0, (r, f, int(bad[0]))
        assert np.array_equal(allpos[r], rep.pos)
    b.close()


@pytest.mark.gpu
def test_gpu_batched_snapshot_restore_continues_the_run(device):
    """A snapshot after 400 legs restored into a fresh batch of the same description continues
    every replica exactly as the original batch does; a snapshot of another configuration is
    refused."""
    from jellyfysh_amd import _lib
    from jellyfysh_amd.batch import DeviceBatch
    ns, lengths, pos = small_batch()
    print("batched snapshot and restore", flush=True)
    a = DeviceBatch(ns, lengths, pos, seed=31, classes=([16384], [1]))
    a.run(400)
    snap = a.snapshot()
    a.run(500)
    b = DeviceBatch(ns, lengths, pos, seed=31, classes=([128, 512, 4096], [1, 2, 4]))
    b.restore(snap)
    b.run(500)
    sa, sb = a.states(), b.states()
    for k in ("active", "direction", "legs", "chain_left", "est", "stats"):
        assert np.array_equal(sa[k], sb[k]), k
    for p1, p2 in zip(a.positions(), b.positions()):
        assert np.array_equal(p1, p2)
    other = DeviceBatch(ns, lengths, pos, seed=32, classes=([16384], [1]))
    with pytest.raises(_lib.EcmcError):
        other.restore(snap)
    with pytest.raises(_lib.EcmcError):
        other.restore(snap[:-8])
    for x in (a, b, other):
        x.close()
