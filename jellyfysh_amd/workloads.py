"""Charged soft spheres on the plain device chain, many replicas in one launch.

The system: N point charges (half +1, half -1) in a periodic cube with the merged-image
Coulomb potential (Ewald alpha = 3.45 / L, Fourier cutoff |n| <= 6, position cutoff 2) plus an
inverse-power-9 repulsion of prefactor 1 between every pair, at beta = 1.

- Coulomb factor: wired as the reference's coulomb_atoms/cell_bounded.ini
  (config_files/2018_JCP_149_064113/coulomb_atoms/cell_bounded.ini): far cells through the
  cell-bounding potential with FP64 Ewald tables (one occupant per cell,
  single_active_cell_occupancy.py:62), nearby cells and the surplus through the
  inverse-power Coulomb bound with thinning.
- Far cells, alternatively: the cell veto of the reference's coulomb_atoms/cell_veto.ini
  (cell_veto_event_handler.py:93-238) over the same estimator bounds (``far="cell_veto"``),
  one sampled target cell per leg in place of one candidate per occupied far cell.
- Repulsion: one more pair tagger over all other units with the exact inversion of
  InversePowerPotential(power=9, prefactor=1) (inverse_power_potential.py:32-179), which has
  no cutoff in the reference.
- Start: uniform random placement with overlaps (minimum-image distance below 0.8) rejected.
  The reference's random input handler places without rejection
  (random_input_handler.py:69-84), so this placement is this package's own, drawn from
  numpy's generator and not from the run's random stream.

``charged_soft_spheres`` returns one system's description; ``charged_soft_sphere_chains``
creates replicas (seeds seed, seed + 1, ...) that ``chain.run_together`` advances in one
launch.
"""
import numpy as np

from .chain import DeviceChain, make_desc

COULOMB_PAIRS = ["coulomb_cell_bounding", "coulomb_nearby", "cell_boundary", "coulomb_surplus", "soft_sphere"]
FAR_TAGS = {"cell_bounding": "coulomb_cell_bounding", "cell_veto": "coulomb_cell_veto"}


def random_placement_without_overlap(n, system_length, min_separation, rng, dim=3, max_attempts=10_000_000):
    """n points placed one after another uniformly in [0, L)^dim; a point closer than
    min_separation (minimum image) to an earlier one is drawn again."""
    L = float(system_length)
    pos = np.empty((n, dim))
    k, attempts = 0, 0
    r2 = float(min_separation) ** 2
    while k < n:
        attempts += 1
        if attempts > max_attempts:
            raise RuntimeError("could not place %d points at separation %g in a box of %g" % (n, min_separation, L))
        x = rng.uniform(0.0, L, dim)
        if k:
            s = pos[:k] - x
            s -= L * np.round(s / L)
            if np.min(np.einsum("ij,ij->i", s, s)) < r2:
                continue
        pos[k] = x
        k += 1
    return pos


def soft_sphere_taggers(alpha=3.45, fourier_cutoff=6, position_cutoff=2, power=9.0, prefactor=1.0, chain_time=1.0,
                        end_of_run_time=1.0e9, far="cell_bounding"):
    """The tagger list (INI sections) of the charged soft spheres; `far` is "cell_bounding" or
    "cell_veto"."""
    if far not in FAR_TAGS:
        raise ValueError("far cells through cell_bounding or cell_veto (got %r)" % (far,))
    mic = [alpha, fourier_cutoff, position_cutoff, 1.0]
    pairs = [FAR_TAGS[far]] + COULOMB_PAIRS[1:]
    eoc = ["end_of_chain"] + pairs
    ipcb = dict(bounding_potential="inverse_power_coulomb_bounding", bounding_params=[1.5837])
    first = dict(tag=FAR_TAGS[far], kind=far, potential="merged_image_coulomb", params=mic,
                 bounding_params=[1.5, 10, 1.0, float("inf")], charge=True, create=pairs, trash=pairs)
    if far == "cell_veto":
        first["number_event_handlers"] = 1
    return [
        first,
        dict(tag="coulomb_nearby", kind="pair_cells", potential="merged_image_coulomb", params=mic, charge=True,
             create=pairs, trash=pairs, **ipcb),
        dict(tag="cell_boundary", kind="cell_boundary", create=pairs, trash=pairs),
        dict(tag="coulomb_surplus", kind="surplus", potential="merged_image_coulomb", params=mic, charge=True,
             create=pairs, trash=pairs, **ipcb),
        dict(tag="soft_sphere", kind="pair_all", potential="inverse_power", params=[power, prefactor], charge=False,
             create=pairs, trash=pairs),
        dict(tag="end_of_chain", kind="end_of_chain", create=eoc, trash=eoc, chain_time=chain_time),
        dict(tag="end_of_run", kind="end_of_run", create=["end_of_run"], trash=eoc + ["end_of_run"],
             end_of_run_time=end_of_run_time),
        dict(tag="start_of_run", kind="start_of_run", create=eoc + ["end_of_run"], trash=["start_of_run"],
             initial_direction_of_motion=0, speed=1.0, initial_active_identifier=0),
    ]


def charged_soft_spheres(n=512, system_length=8.0, seed=0, min_separation=0.8, cells_per_side=8, beta=1.0,
                         chain_time=1.0, far="cell_bounding"):
    """One replica: dict with dim, n, system_length, beta, cells_per_side, seed, taggers,
    positions (n x 3) and charges (+1 for even identifiers, -1 for odd)."""
    if n % 2:
        raise ValueError("an even number of charges")
    rng = np.random.default_rng(seed)
    pos = random_placement_without_overlap(n, system_length, min_separation, rng)
    charges = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return dict(dim=3, n=n, system_length=float(system_length), beta=float(beta), cells_per_side=int(cells_per_side),
                seed=int(seed), taggers=soft_sphere_taggers(chain_time=chain_time, far=far), positions=pos, charges=charges,
                max_occupants=1)


def make_chain(system):
    desc = make_desc(system["dim"], system["n"], system["system_length"], system["beta"], system["cells_per_side"],
                     system["taggers"], system["seed"], scheduler="heap", max_occupants=system["max_occupants"])
    return DeviceChain(desc, system["positions"], system["charges"])


def charged_soft_sphere_chains(replicas, seed=0, **kw):
    """Device chains of `replicas` independent charged-soft-sphere systems (seeds seed + i)."""
    return [make_chain(charged_soft_spheres(seed=seed + i, **kw)) for i in range(replicas)]
