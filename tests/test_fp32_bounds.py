"""The FP32 bounds of the batched kernel rest on two measured facts about the device's
FP32 math: log2f's error relative to max(|log2 x|, 1) and exp2f's relative error.  These
scans run every float of the domains the bounds use and check the errors against the
margins the kernel applies (LOG2F_ERR = 2^-20 for the draw's dU bound; ETA_POW = 2^-14
around exp2f(2/p log2f(q)), where an exponent of magnitude <= 128 turns a log2f error of
2^-21 into a relative error below 2^-19)."""
import struct

import numpy as np
import pytest


def _scan(device, fn, lo, hi):
    import torch
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    device.check(device.lib().ecmc_fp32_math_error(fn, lo, hi, device.dptr(out), None))
    torch.cuda.synchronize()
    return struct.unpack("<d", struct.pack("<q", int(out.item())))[0]


@pytest.mark.gpu
def test_log2f_and_exp2f_errors_stay_inside_the_bound_margins(device):
    print("scanning log2f over every positive float, exp2f over [-126, 128)", flush=True)
    e_log = _scan(device, 0, 0x00000001, 0x7F800000)
    e_exp = max(_scan(device, 1, 0x80000000, 0xC2FC0001), _scan(device, 1, 0x00000000, 0x43000000))
    print("log2f max error / max(|log2 x|, 1) 2^%.2f, exp2f max relative error 2^%.2f"
          % (__import__("math").log2(e_log) if e_log else -999, __import__("math").log2(e_exp) if e_exp else -999),
          flush=True)
    assert e_log <= 2.0 ** -21
    assert e_exp <= 2.0 ** -21


def _bound_rows(rng, n, dim, L, near):
    """Separations over the box, half of them placed near the edge cases of the inversion:
    the target level with the active unit along the motion, on the line of motion, at the
    cutoff sphere."""
    sep = rng.uniform(-L / 2, L / 2, (n, dim))
    d = rng.integers(0, dim, n).astype(np.int32)
    k = rng.integers(0, 4, n)
    rows = np.arange(n)
    m = near & (k == 0)
    sep[rows[m], d[m]] = rng.normal(0, 1e-6, m.sum())                       # level along the motion
    m = near & (k == 1)
    for q in range(dim):
        sel = m & (d != q)
        sep[rows[sel], q] *= 1e-7                                            # on the line of motion
    m = near & (k == 2)
    sep[m] *= (2.5 / np.linalg.norm(sep[m], axis=1) * (1 + rng.normal(0, 1e-7, m.sum())))[:, None]
    return sep, d


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [2, 3])
def test_gpu_pair_time_brackets_hold_the_exact_displacement(device, dim):
    """ecmc_pair_time_bounds (the chain kernel's FP32 candidate filter) against the exact
    displacement of ecmc_pair_displacement on the same rows, the potential change drawn from the
    same random words: lb <= t <= ub on every row, lb = +inf only where t = +inf, ub finite only
    where t is (where lb = 0, t may also be a rounding-level negative number, the FP64 form of an
    event right at the start); powers 1, 6, 9, 12, both signs, charges, plain and truncated (rc
    2.5), two speeds.
    Also the brackets are tight: where an event is certain, the median of (ub - lb) / t stays
    below 5e-3."""
    import ctypes
    import torch
    rng = np.random.default_rng(40 + dim)
    n, L = 200000, 8.0
    dev = lambda a, dt: torch.tensor(a, dtype=dt, device="cuda").contiguous()
    checked = certain = 0
    for potential, trunc in ((device.POT_INVERSE_POWER, False), (device.POT_INVERSE_POWER_TRUNC, True)):
        for power in (1, 6, 9, 12):
            for prefactor in (1.0, -1.0, 4.0):
                print("brackets: dim %d, power %d, prefactor %g, truncated %s" % (dim, power, prefactor, trunc),
                      flush=True)
                near = rng.random(n) < 0.5
                sep, d = _bound_rows(rng, n, dim, L, near)
                speed = np.where(rng.random(n) < 0.5, 1.0, 0.7)
                c1 = rng.choice([1.0, -1.0, 0.5], n)
                c2 = rng.choice([1.0, -1.0], n)
                w = rng.integers(0, 2 ** 32, (n, 2), dtype=np.uint64).astype(np.uint32)
                beta = float(rng.uniform(0.5, 2.0))
                u = ((w[:, 0] >> 5).astype(np.float64) * 67108864.0 + (w[:, 1] >> 6)) * (1.0 / 9007199254740992.0)
                dU = -np.log(1.0 - u) / beta
                par = np.array([power, prefactor, 2.5, 0.0])
                pptr = par.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                t_dir, t_speed, t_sep = dev(d, torch.int32), dev(speed, torch.float64), dev(sep, torch.float64)
                t_c1, t_c2 = dev(c1, torch.float64), dev(c2, torch.float64)
                t_word = dev(w[:, 0].view(np.int32), torch.int32)
                disp = torch.empty(n, dtype=torch.float64, device="cuda")
                lb = torch.empty(n, dtype=torch.float32, device="cuda")
                ub = torch.empty(n, dtype=torch.float32, device="cuda")
                device.check(device.lib().ecmc_pair_displacement(potential, pptr, dim, n, device.dptr(t_dir),
                                                                 device.dptr(t_speed), device.dptr(t_sep),
                                                                 device.dptr(t_c1), device.dptr(t_c2),
                                                                 device.dptr(dev(dU, torch.float64)), L,
                                                                 device.dptr(disp), None))
                device.check(device.lib().ecmc_pair_time_bounds(potential, pptr, dim, n, device.dptr(t_dir),
                                                                device.dptr(t_speed), device.dptr(t_sep),
                                                                device.dptr(t_c1), device.dptr(t_c2),
                                                                device.dptr(t_word), beta, device.dptr(lb),
                                                                device.dptr(ub), None))
                torch.cuda.synchronize()
                t, lo, hi = disp.cpu().numpy(), lb.cpu().numpy().astype(np.float64), ub.cpu().numpy().astype(np.float64)
                ok = ~np.isnan(t)
                t, lo, hi = t[ok], lo[ok], hi[ok]
                # lb = 0 also covers the rounding-level negative displacements of an event at the start
                low_ok = (lo <= t) | ((lo == 0) & (t < 0) & (t > -1e-9))
                assert np.all(low_ok), (power, prefactor, trunc, int(np.sum(~low_ok)))
                assert np.all(t <= hi), (power, prefactor, trunc, int(np.sum(t > hi)))
                assert np.all(np.isinf(t[np.isinf(lo)]))
                fin = np.isfinite(hi)
                if fin.sum() >= 100:
                    gap = (hi[fin] - lo[fin]) / np.maximum(t[fin], 1e-300)
                    assert np.median(gap) < 5e-3, (power, prefactor, trunc, np.median(gap))
                checked += int(ok.sum())
                certain += int(fin.sum())
    assert checked > 20 * n and certain > n


@pytest.mark.gpu
def test_gpu_coulomb_bound_time_brackets_hold_the_exact_displacement(device):
    """ecmc_pair_time_bounds for the inverse-power Coulomb bound (the chain kernel's filter of the
    merged-image Coulomb candidates) against ecmc_pair_displacement's exact inversion on the same
    rows and the same random words: lb <= t <= ub on every row, lb always finite (an event always
    comes).  Rows over the whole box and near the special cases: the target on the line of motion
    (rho -> 0), at the corner (rho -> half sqrt 2), level with the active unit along the motion and
    half a box away; both signs, two box lengths, two speeds.  The brackets must resolve the
    windings: where the exact count floor(dU / per) is at least 2, lb > 0."""
    import ctypes
    import torch
    rng = np.random.default_rng(77)
    n = 200000
    dev = lambda a, dt: torch.tensor(a, dtype=dt, device="cuda").contiguous()
    checked = resolved = 0
    for L in (8.0, 21.5):
        for prefactor in (1.5837, 0.4):
            print("Coulomb-bound brackets: L %g, prefactor %g" % (L, prefactor), flush=True)
            h = L / 2
            sep = rng.uniform(-h, h, (n, 3))
            d = rng.integers(0, 3, n).astype(np.int32)
            rows = np.arange(n)
            k = rng.integers(0, 6, n)
            for q in range(3):
                sel = (k == 0) & (d != q)
                sep[rows[sel], q] *= 1e-7                                   # on the line of motion
                sel = (k == 1) & (d != q)
                sep[rows[sel], q] = np.sign(sep[rows[sel], q]) * h * (1 - rng.uniform(0, 1e-9, sel.sum()))  # corner
            sel = k == 2
            sep[rows[sel], d[sel]] = rng.normal(0, 1e-9, sel.sum())          # level along the motion
            sel = k == 3
            sep[rows[sel], d[sel]] = -h * (1 - rng.uniform(0, 1e-12, sel.sum()))   # half a box away
            speed = np.where(rng.random(n) < 0.5, 1.0, 0.7)
            c1 = rng.choice([1.0, -1.0, 0.5, -0.5], n)
            c2 = rng.choice([1.0, -1.0, 0.5], n)
            w = rng.integers(0, 2 ** 32, (n, 2), dtype=np.uint64).astype(np.uint32)
            beta = float(rng.uniform(0.5, 2.0))
            u = ((w[:, 0] >> 5).astype(np.float64) * 67108864.0 + (w[:, 1] >> 6)) * (1.0 / 9007199254740992.0)
            dU = -np.log(1.0 - u) / beta
            par = np.array([prefactor, L, 0.0, 0.0])
            pptr = par.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            t_dir, t_speed, t_sep = dev(d, torch.int32), dev(speed, torch.float64), dev(sep, torch.float64)
            t_c1, t_c2 = dev(c1, torch.float64), dev(c2, torch.float64)
            t_word = dev(w[:, 0].view(np.int32), torch.int32)
            disp = torch.empty(n, dtype=torch.float64, device="cuda")
            lb = torch.empty(n, dtype=torch.float32, device="cuda")
            ub = torch.empty(n, dtype=torch.float32, device="cuda")
            device.check(device.lib().ecmc_pair_displacement(device.POT_INVERSE_POWER_COULOMB_BOUND, pptr, 3, n,
                                                             device.dptr(t_dir), device.dptr(t_speed), device.dptr(t_sep),
                                                             device.dptr(t_c1), device.dptr(t_c2),
                                                             device.dptr(dev(dU, torch.float64)), L, device.dptr(disp), None))
            device.check(device.lib().ecmc_pair_time_bounds(device.POT_INVERSE_POWER_COULOMB_BOUND, pptr, 3, n,
                                                            device.dptr(t_dir), device.dptr(t_speed), device.dptr(t_sep),
                                                            device.dptr(t_c1), device.dptr(t_c2), device.dptr(t_word), beta,
                                                            device.dptr(lb), device.dptr(ub), None))
            torch.cuda.synchronize()
            t, lo, hi = disp.cpu().numpy(), lb.cpu().numpy().astype(np.float64), ub.cpu().numpy().astype(np.float64)
            assert np.all(np.isfinite(lo))
            ok = ~np.isnan(t)
            assert ok.sum() > 0.99 * n
            t, lo, hi = t[ok], lo[ok], hi[ok]
            assert np.all(lo <= t), (L, prefactor, int(np.sum(lo > t)))
            assert np.all(t <= hi), (L, prefactor, int(np.sum(t > hi)))
            # the windings: floor(dU / per) >= 2 must give lb > 0
            rho = np.sqrt(np.sum(sep ** 2, axis=1) - sep[rows, d] ** 2)[ok]
            pp = np.abs(prefactor * c1 * c2)[ok]
            with np.errstate(divide="ignore"):
                per = pp * (1 / rho - 1 / np.sqrt(h * h + rho * rho))
            wind = np.floor(dU[ok] / per) >= 2
            assert np.all(lo[wind] > 0), (L, prefactor, int(np.sum(lo[wind] == 0)))
            checked += int(ok.sum())
            resolved += int(wind.sum())
    assert checked > 3 * n and resolved > n
