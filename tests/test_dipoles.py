"""Dipoles from the INI front end: composite point objects with factor type maps, the Coulomb
factor through the inverse-power Coulomb bound with thinning.

- jellyfysh_amd/configs/dipoles_atom_factors.ini (the reference's dipoles/atom_factors.ini):
  the device against the oracle event for event, and 256 replicas sampling the separations of
  leaf units of different dipoles against the reference's reversible-Monte-Carlo data
  (output/2018_JCP_149_064113/dipoles/ReferenceDataDipoles_13.dat and _14.dat, copied by
  tests/golden/copy_dipoles_cdf.py), as the reference's plotting script compares its run.
- jellyfysh_amd/configs/flexible_dipoles.ini (500 dipoles, +-0.5 charges, bond r0 = 1, k = 400,
  L = 21.544): the device against the oracle, 64 replicas in one launch against replicas run
  alone, and dU/d(k, r0, q) of the energy's backward kernels against finite differences.
"""
import os

import numpy as np
import pytest

import systems as S

CONFIGS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "jellyfysh_amd", "configs")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _config(name):
    from jellyfysh_amd.config import RunConfig
    return RunConfig.from_file(os.path.join(CONFIGS, name + ".ini"))


def _oracle(oracle, cfg, seed, math_mode, ties=0):
    """ties: the oracle's rule for events of equal time; the device's comparisons take
    first push (oracle.TIES_FIRST_PUSH), since sampling times and chain ends coincide exactly in
    these runs (sampling every 50, chains of 2)."""
    pos, q, words = cfg.initial_state(seed)
    oc = S.oracle_config_from_wiring(cfg.dim, len(pos), cfg.system_length, cfg.beta, 1, cfg.taggers, seed,
                                     cfg.scheduler, math_mode, cfg.max_occupants, leaves_per_root=cfg.leaves_per_root,
                                     ties=ties)
    oc.rng_skip = words
    return oracle.Engine(oc, pos, q)


def _device(cfg, seed):
    from jellyfysh_amd.chain import DeviceChain
    pos, q, words = cfg.initial_state(seed)
    return DeviceChain(cfg.desc(len(pos), seed, rng_skip=words), pos, q)


def test_dipole_inis_read_as_factor_maps():
    a = _config("dipoles_atom_factors")
    assert a.leaves_per_root == 2
    kinds = {t["tag"]: (t["kind"], t.get("factor_map")) for t in a.taggers}
    assert kinds["coulomb"] == ("factor_map", {"local": False, "targets": [[2, 3], [2, 3]]})
    assert kinds["harmonic"] == ("factor_map", {"local": True, "targets": [[1], [0]]})
    assert kinds["repulsive"] == ("factor_map", {"local": False, "targets": [[3], [2]]})
    pos, q, words = a.initial_state(5)
    assert pos.shape == (4, 3) and list(q) == [1.0, -1.0, 1.0, -1.0]
    f = _config("flexible_dipoles")
    pos, q, words = f.initial_state(5)
    assert pos.shape == (1000, 3) and abs(f.system_length ** 3 * 0.05 - 500) < 0.05
    s = pos[1::2] - pos[0::2]
    s -= f.system_length * np.round(s / f.system_length)
    d = np.sqrt((s * s).sum(1))
    assert np.all((d >= 0.9 - 1e-12) & (d <= 1.1 + 1e-12))
    assert [t["tag"] for t in f.taggers][:3] == ["coulomb", "harmonic", "repulsive"]
    assert f.taggers[0]["number_event_handlers"] == 998


def test_oracle_runs_the_flexible_dipoles(oracle):
    e = _oracle(oracle, _config("flexible_dipoles"), 3, oracle.MATH_LIBM)
    log = e.run(400)
    assert len(log) == 400
    assert abs(log["ncand"][log["tagger"] != 6].mean() - 1997) < 5      # 998 + 1 + 998 per leg
    counts = np.bincount(log["tagger"], minlength=7)
    assert counts[0] > 0 and counts[1] > 0 and counts[2] > 0


def _compare(got, want, times_rel=1e-12):
    assert len(got) == len(want)
    for f in ("tagger", "flags", "active", "target", "ncand"):
        bad = np.flatnonzero(got[f] != want[f])
        assert bad.size == 0, "field %s differs first at leg %d" % (f, bad[0])
    t_got, t_want = got["tq"] + got["tr"], want["tq"] + want["tr"]
    assert np.max(np.abs(t_got - t_want) / np.maximum(np.abs(t_want), 1.0)) < times_rel


@pytest.mark.gpu
def test_gpu_dipole_atom_factors_match_oracle(device, oracle):
    cfg = _config("dipoles_atom_factors")
    legs = 6000
    print("two dipoles (atom_factors.ini), %d legs" % legs, flush=True)
    chain = _device(cfg, 11)
    got = chain.run(legs)
    e = _oracle(oracle, cfg, 11, oracle.MATH_CR, oracle.TIES_FIRST_PUSH)
    want = e.run(legs)
    _compare(got, want)
    assert chain.rng_words() == e.rng_words()
    assert np.bincount(got["tagger"], minlength=7)[0] > 50           # Coulomb events commit


@pytest.mark.gpu
def test_gpu_dipoles_sample_the_reference_distribution(device):
    """256 replicas of dipoles_atom_factors.ini (seeds 0-255) advanced together, each stopping at its
    sampling events: after 20 samples of burn-in per replica, the separations of the leaf units
    of different dipoles with equal leaf numbers (the reference's _13 file) and with different
    ones (_14) against the reference's reversible-Monte-Carlo curves.  About 2 x 195 000
    correlated samples per curve; the bound 0.015 on the Kolmogorov distance leaves room for
    their correlation."""
    from jellyfysh_amd.chain import run_together
    cfg = _config("dipoles_atom_factors")
    samp = [i for i, t in enumerate(cfg.taggers) if t["kind"] == "sampling"][0]
    chains = []
    for seed in range(256):
        c = _device(cfg, seed)
        c.set_stop_mask(1 << samp)
        chains.append(c)
    L = cfg.system_length
    s13, s14 = [], []
    print("256 replicas of two dipoles, 400 sampling stops", flush=True)
    for it in range(400):
        n, _ = run_together(chains, 200000)
        assert np.all((n > 0) & (n < 200000))
        if it < 20:
            continue
        for c in chains:
            p = c.positions()
            for a, b, out in ((0, 2, s13), (1, 3, s13), (0, 3, s14), (1, 2, s14)):
                s = np.mod((p[b] - p[a]) + L / 2, L) - L / 2
                out.append(np.sqrt(np.sum(s * s)))
        if it % 100 == 0:
            print("stop %d" % it, flush=True)
    for c in chains:
        c.close()
    for part, seps in (("13", s13), ("14", s14)):
        ref = np.loadtxt(os.path.join(GOLDEN, "dipoles_reference_cdf_%s.dat" % part))
        seps = np.sort(np.array(seps))
        F = np.searchsorted(seps, ref[:, 0], side="right") / len(seps)
        ks = np.max(np.abs(F - ref[:, 1]))
        print("_%s: samples %d, Kolmogorov distance to the reference data %.4f, median %.4f" %
              (part, len(seps), ks, np.median(seps)), flush=True)
        assert ks < 0.015


@pytest.mark.gpu
def test_gpu_flexible_dipoles_match_oracle(device, oracle):
    cfg = _config("flexible_dipoles")
    legs = 1200
    print("500 flexible dipoles, %d legs" % legs, flush=True)
    chain = _device(cfg, 3)
    got = chain.run(legs)
    e = _oracle(oracle, cfg, 3, oracle.MATH_CR, oracle.TIES_FIRST_PUSH)
    want = e.run(legs)
    _compare(got, want)
    assert chain.rng_words() == e.rng_words()
    dpos = np.abs(chain.positions() - e.positions())
    assert np.max(np.minimum(dpos, cfg.system_length - dpos)) < 1e-9


@pytest.mark.gpu
def test_gpu_flexible_dipole_replicas_together(device):
    """64 replicas in one launch (the flexible-dipole workload), two of them replayed alone."""
    from jellyfysh_amd.chain import run_together
    cfg = _config("flexible_dipoles")
    legs = 300
    print("64 flexible-dipole replicas, %d legs" % legs, flush=True)
    chains = [_device(cfg, 100 + i) for i in range(64)]
    n, ms = run_together(chains, legs)
    assert list(n) == [legs] * 64 and ms > 0
    for i in (0, 63):
        alone = _device(cfg, 100 + i)
        assert alone.run(legs, log=False) == legs
        assert np.array_equal(alone.positions(), chains[i].positions())
        assert np.array_equal(alone.estimators(), chains[i].estimators())
        assert alone.rng_words() == chains[i].rng_words()
    for c in chains:
        c.close()


@pytest.mark.gpu
def test_gpu_flexible_dipole_parameter_derivatives(device):
    """dU/dk and dU/dr0 of the bond (displaced even power 2, within a dipole), dU/dq of the charge
    magnitude (Coulomb between dipoles, charges +-q) and dU/dk12 of the repulsion's prefactor
    from the hand-written backward kernels, against central finite differences of the forward
    energy, at a configuration of the flexible-dipole workload."""
    import torch
    from jellyfysh_amd.energy import pair_energy
    cfg = _config("flexible_dipoles")
    pos, q, _ = cfg.initial_state(4)
    L = cfg.system_length
    x = torch.tensor(pos, dtype=torch.float64, device="cuda")
    mol = torch.arange(len(pos), device="cuda", dtype=torch.int32) // 2
    sign = torch.tensor(np.sign(q), dtype=torch.float64, device="cuda")

    def energies(k, r0, qq, k12):
        bond = pair_energy(x, torch.stack([r0, torch.tensor(2.0, dtype=torch.float64, device="cuda"), k / 2]),
                           "displaced_even_power", L, molecule=mol, scope="intra")
        coul = pair_energy(x, [3.45, 6.0, 2.0, 1.0], "merged_image_coulomb", L, charges=qq * sign, molecule=mol,
                           scope="inter")
        rep = pair_energy(x, torch.stack([torch.tensor(12.0, dtype=torch.float64, device="cuda"), k12]),
                          "inverse_power", L, molecule=mol, scope="inter")
        return bond, coul, rep

    theta = [torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True) for v in (400.0, 1.0, 0.5, 1.0)]
    sum(energies(*theta)).backward()
    grads = [float(t.grad) for t in theta]
    # the finite difference of the one term each parameter enters (the bond for k and r0, the
    # Coulomb term for q, the repulsion for k12): the repulsion of the random start reaches
    # ~1e8, whose rounding would swamp a difference of the summed energy
    term = (0, 0, 1, 2)
    for i, h in enumerate((1e-3, 1e-6, 1e-6, 1e-6)):
        args_p = [float(v) for v in theta]
        args_m = [float(v) for v in theta]
        args_p[i] += h
        args_m[i] -= h
        up = float(energies(*[torch.tensor(v, dtype=torch.float64, device="cuda") for v in args_p])[term[i]])
        um = float(energies(*[torch.tensor(v, dtype=torch.float64, device="cuda") for v in args_m])[term[i]])
        fd = (up - um) / (2 * h)
        assert abs(grads[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, grads[i], fd)


@pytest.mark.gpu
def test_gpu_hard_disk_dipoles_sample_the_reference_polarization(device):
    """256 replicas of hard_disk_dipoles_cells.ini (the reference's 81 hard-disk dipoles from its
    .pdb start, seeds 0-255) advanced together against the reference's Newtonian-ECMC
    polarization curves (output/hard_disk_dipoles/ReferenceDataPx/Py_81Dipoles_NewtonianECMC.dat,
    copied by tests/golden/copy_hard_disk_dipoles_cdf.py).  The polarization relaxes slowly: its
    autocorrelation falls to 0.4 after about 28 000 time units (18 legs per time unit), so each
    replica first runs 1.5 M legs (about 83 000 time units) without stopping, then samples four
    times, each at the first sampling event (every 10.01) after another 0.75 M legs; 1024 samples
    of about 500 independent ones.  The Kolmogorov distance to the reference curves must stay
    below 0.1: the same protocol on the oracle gives about 0.04, and the reference's own run of
    15 M time units holds a few hundred independent samples."""
    from jellyfysh_amd.chain import DeviceChain, run_together
    from jellyfysh_amd.output import PolarizationOutputHandler
    cfg = _config("hard_disk_dipoles_cells")
    samp = [i for i, t in enumerate(cfg.taggers) if t["kind"] == "sampling"][0]
    pos, q, _ = cfg.initial_state(0)
    chains = [DeviceChain(cfg.desc(len(pos), seed), pos) for seed in range(256)]
    pol = PolarizationOutputHandler.polarization     # the output handler's sum, without its file
    burn, spacing, n_samples, piece = 1500000, 750000, 4, 250000

    def advance(legs):
        for c in chains:
            c.set_stop_mask(0)
        for _ in range(legs // piece):
            n, ms = run_together(chains, piece)
            assert np.all(n == piece)
            print("  %d legs per replica in %.0f ms" % (piece, ms), flush=True)

    print("256 replicas of 81 hard-disk dipoles: burn-in %d legs" % burn, flush=True)
    advance(burn)
    samples = []
    for k in range(n_samples):
        advance(spacing)
        for c in chains:
            c.set_stop_mask(1 << samp)
        n, _ = run_together(chains, 100000)
        assert np.all((n > 0) & (n < 100000))              # every replica stopped at a sampling event
        samples += [pol(_PolarizationView(), c.positions(), cfg.system_length, q) for c in chains]
        print("sample %d taken" % k, flush=True)
    for c in chains:
        c.close()
    samples = np.array(samples)
    for k, part in enumerate(("px", "py")):
        ref = np.loadtxt(os.path.join(GOLDEN, "hard_disk_dipoles_reference_cdf_%s.dat" % part))
        v = np.sort(samples[:, k])
        F = np.searchsorted(v, ref[:, 0], side="right") / len(v)
        ks = np.max(np.abs(F - ref[:, 1]))
        print("%s: samples %d, Kolmogorov distance to the reference data %.4f, mean %.3f, std %.3f" %
              (part, len(v), ks, v.mean(), v.std()), flush=True)
        assert ks < 0.1


class _PolarizationView:
    """Stands for a PolarizationOutputHandler of two leaves per object (no output file)."""
    leaves_per_root = 2
