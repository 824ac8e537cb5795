"""The truncated-and-shifted inverse-power inversion (ora_ipt_displacement, the function every
event of the batched workload goes through; it has no reference counterpart) pinned without
restating its formula:

- energy: at the returned displacement the potential, evaluated in 60-digit arithmetic,
  has risen by exactly the drawn potential change from the reference point of the rise (the
  start, the closest approach for an attractive factor ahead, or the cutoff sphere for a
  repulsive factor entering it), within the rounding of the returned displacement
  propagated through the potential's slope;
- brute force: along a fine grid of the straight path, the largest energy rise over any
  earlier minimum stays below the draw where the inversion returns +inf, and crosses it at
  the returned displacement where it is finite;
- the limit rc -> infinity: the reference's own inverse-power unit-test vectors
  (unittests/test_potential/test_inverse_power_potential.py) are reproduced by the truncated
  form with a cutoff far beyond every separation.
"""
import math

import numpy as np
import pytest

import refapi

mp = pytest.importorskip("mpmath")
mp.mp.dps = 60

EPS = 2.0 ** -53


def _ipt(oracle, s, d, power, k, rc, dU, mode):
    L = oracle.lib()
    return L.ora_ipt_displacement(3, d, 1.0, oracle._dp(oracle.vec(s)), 1.0, 1.0, power, k, rc, dU, mode)


def _u(k, power, r2):
    """Unshifted k / r^p in 60 digits from a squared distance (mpf)."""
    return mp.mpf(k) / r2 ** (mp.mpf(power) / 2)


def _case(rng):
    power = float(rng.choice([6.0, 12.0]))
    k = float(rng.choice([4.0, -4.0]) * rng.uniform(0.5, 2.0))
    rc = float(rng.uniform(1.5, 3.0))
    d = int(rng.integers(3))
    s = rng.uniform(-1.2 * rc, 1.2 * rc, 3)
    dU = float(rng.exponential(1.5))
    return power, k, rc, d, s, dU


def _reference_point(s, d, k, rc):
    """Squared distance at which the rise starts: the start, or (attractive, unit ahead) the
    closest approach; for a repulsive factor outside the cutoff the rise starts on the sphere."""
    s = [mp.mpf(float(x)) for x in s]
    if k < 0 and s[d] > 0:
        s[d] = mp.mpf(0)
    r2 = sum(x * x for x in s)
    return min(r2, mp.mpf(rc) ** 2)


@pytest.mark.parametrize("mode", [0, 1])
def test_energy_rises_by_the_draw_at_the_returned_displacement(oracle, mode):
    rng = np.random.default_rng(20 + mode)
    checked = finite = 0
    worst = 0.0
    while checked < 100_000:
        power, k, rc, d, s, dU = _case(rng)
        dt = _ipt(oracle, s, d, power, k, rc, dU, mode)
        checked += 1
        if not math.isfinite(dt):
            continue
        finite += 1
        # an event right at the start (a tiny draw inside the cutoff) can come out a rounding
        # below zero, as the reference's own InversePowerPotential inversion does: s_d - sqrt(a)
        # with sqrt(a) within an ulp of s_d
        assert dt >= -16 * EPS * float(max(abs(x) for x in s))
        sm = [mp.mpf(float(x)) for x in s]
        sm[d] -= mp.mpf(dt)
        r2e = sum(x * x for x in sm)
        assert r2e <= mp.mpf(rc) ** 2 * (1 + 1e-12)          # the event lies inside the cutoff
        u0 = _u(k, power, _reference_point(s, d, k, rc))
        ue = _u(k, power, r2e)
        rise = ue - u0
        # slope of U along the path at the event, times the rounding of dt and of the
        # separation components it was formed from
        slope = abs(mp.mpf(power) * k * sm[d] / r2e ** (mp.mpf(power) / 2 + 1))
        dx = 16 * EPS * (abs(dt) + float(max(abs(x) for x in s)) + math.sqrt(float(r2e)))
        tol = slope * dx + 16 * EPS * power * (abs(u0) + abs(ue))
        err = abs(rise - dU)
        worst = max(worst, float(err / tol))
        assert err <= tol, (power, k, rc, d, s.tolist(), dU, dt, float(rise))
    assert finite > 8_000
    print("finite events %d of %d, worst error / bound %.3g" % (finite, checked, worst))


def _path_rise(s, d, power, k, rc, t):
    """Energy of the shifted potential along the path (numpy, float64) at displacements t."""
    sd = s[d] - t
    perp2 = sum(s[q] ** 2 for q in range(3) if q != d)
    r2 = perp2 + sd * sd
    u = np.where(r2 < rc * rc, k / r2 ** (power / 2) - k / rc ** power, 0.0)
    return u


def test_brute_force_along_the_path(oracle):
    """+inf only where no point of the path rises by the draw over an earlier minimum, and a
    finite displacement where the rise first reaches the draw (grid of 20001 points over the
    path through the cutoff sphere)."""
    rng = np.random.default_rng(5)
    n_inf = n_fin = 0
    for _ in range(8000):
        power, k, rc, d, s, dU = _case(rng)
        # keep clear of the potential's singular core, where a grid cannot resolve the rise
        perp2 = sum(s[q] ** 2 for q in range(3) if q != d)
        if perp2 < 0.6 ** 2:
            continue
        dt = _ipt(oracle, s, d, power, k, rc, dU, 1)
        tmax = max(s[d] + rc + 1.0, 1.0)
        # the grid plus the closest approach, where an attractive factor's rise starts
        t = np.union1d(np.linspace(0.0, tmax, 20001), [min(max(s[d], 0.0), tmax)])
        u = _path_rise(s, d, power, k, rc, t)
        rise = u - np.minimum.accumulate(u)
        if math.isinf(dt):
            n_inf += 1
            assert rise.max() < dU * (1 + 1e-9) + 1e-12, (power, k, rc, d, s.tolist(), dU, rise.max())
        else:
            n_fin += 1
            # the rise at dt over the running minimum of the grid points before it
            ue = _path_rise(s, d, power, k, rc, np.array([dt]))[0]
            umin = min(float(u[t < dt].min()) if np.any(t < dt) else ue, ue)
            assert abs((ue - umin) - dU) <= 1e-6 * max(1.0, dU), (power, k, rc, d, s.tolist(), dU, dt, ue - umin)
            before = rise[t < dt - (t[1] - t[0])]
            assert before.size == 0 or before.max() <= dU * (1 + 1e-6)
            assert -1e-15 <= dt <= tmax
    assert n_inf > 300 and n_fin > 400


def test_large_cutoff_reproduces_the_reference_inverse_power_vectors(oracle):
    """InversePowerPotential's displacement vectors of the reference's unit tests, through the
    truncated form with rc = 1e150: the shift k / rc^p vanishes and every separation lies inside."""
    vecs = [v for v in refapi_vectors() if v["potential"]["cls"] == "InversePowerPotential"
            and v["method"] == "displacement" and refapi.axis_velocity(v["args"][0]) is not None]
    assert len(vecs) >= 10
    L = oracle.lib()
    for v in vecs:
        par = refapi.ctor_params(v["potential"])
        vel, s, c1, c2, dU = refapi.split_args(v)
        d, speed = refapi.axis_velocity(vel)
        got = L.ora_ipt_displacement(len(s), d, speed, oracle._dp(oracle.vec(s)), c1, c2, par[0], par[1], 1e150, dU, 0)
        assert refapi.matches(got, v), (v["line"], got, v["expected"])


def refapi_vectors():
    import json
    import os
    return json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "reference_unit_vectors.json")))
