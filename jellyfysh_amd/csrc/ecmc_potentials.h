// ecmc_potentials.h -- pair potentials of the ECMC hot path, FP64, device side.
//
// Each function keeps the operation order of the reference's Python/C expression it
// stands in for (cited per function), because a different order rounds differently and
// the event sequence of a chain depends on every bit.  `pw`/`lg` are the correctly rounded
// pow/log of ecmc_math.h.
#pragma once
#include "ecmc_math.h"

namespace ecmc {

enum PotentialKind {
    POT_HARD_SPHERE = 0,          // par: radius
    POT_INVERSE_POWER = 1,        // par: power, prefactor
    POT_LENNARD_JONES = 2,        // par: prefactor, sigma
    POT_INVERSE_POWER_TRUNC = 3,  // par: power, prefactor, cutoff
    POT_MERGED_IMAGE_COULOMB = 4, // par: alpha, fourier_cutoff, position_cutoff, prefactor
    POT_INVERSE_POWER_COULOMB_BOUND = 5,  // par: prefactor
    POT_DISPLACED_EVEN_POWER = 6, // par: equilibrium_separation, power, prefactor
    POT_CELL_BOUNDING = 7,        // cell-bounding tables (no pair evaluation)
    POT_HARD_DIPOLE = 8           // par: minimum_separation, maximum_separation
};

// base/vectors.py:46-60 -- sum() starts from the integer 0.
template <int D>
ECMC_FN double norm_sq(const double *s) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++) acc = acc + s[i] * s[i];
    return acc;
}
// sum([v ** 2 for v in others]) with v ** 2 == v * v (pow(v, 2.0) is exact)
template <int D>
ECMC_FN double others_sq(const double *s, int dir) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++)
        if (i != dir) acc = acc + s[i] * s[i];
    return acc;
}

// HardSpherePotential.displacement (potential/hard_sphere_potential.py:65-99) for a velocity
// with the single positive component `speed` along `dir`.
template <int D>
ECMC_FN double hs_displacement(int dir, double speed, const double *s, double radius) {
    double diam2 = 4.0 * radius * radius;
    double v[D];
#pragma unroll
    for (int i = 0; i < D; i++) v[i] = i == dir ? speed : 0.0;
    double v2 = norm_sq<D>(v);
    double s2 = norm_sq<D>(s);
    double vs = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++) vs = vs + v[i] * s[i];
    double srt = vs * vs - v2 * (s2 - diam2);
    return (srt >= 0.0 && vs >= 0.0) ? (vs - sqrt(srt)) / v2 : INFINITY;
}

// HardDipolePotential.displacement (hard_dipole_potential.py:66-110): the time until the
// separation reaches the minimum (approaching, if it can) or the maximum separation; NaN when
// the separation lies outside [minimum, maximum] by more than 1e-13 (the reference asserts).
template <int D>
ECMC_FN double hd_displacement(int dir, double speed, const double *s, double rmin, double rmax) {
    const double min2 = rmin * rmin, max2 = rmax * rmax;
    double v[D];
#pragma unroll
    for (int i = 0; i < D; i++) v[i] = i == dir ? speed : 0.0;
    const double v2 = norm_sq<D>(v);
    const double s2 = norm_sq<D>(s);
    if (!(s2 - min2 > -1.0e-13) || !(max2 - s2 > -1.0e-13) || !(v2 > 0.0)) return NAN;
    double vs = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++) vs = vs + v[i] * s[i];
    if (vs >= 0.0) {
        const double m = vs * vs - v2 * (s2 - min2);
        if (m >= 0.0) return (vs - sqrt(m)) / v2;
    }
    const double M = vs * vs - v2 * (s2 - max2);
    return M >= 0.0 ? (vs + sqrt(M)) / v2 : NAN;
}

// InversePowerPotential.potential (inverse_power_potential.py:127-143)
// Correctly rounded r2 ** (power / 2) and x ** (2 / power) of the inverse-power inversions:
// the double-double forms inline for the powers 6 and 12 (pow36, root36), else cr_pow.
// Odd powers take the half-integer and 2/p fast paths of ecmc_math.h inline, with their
// rounding tests.
ECMC_FN double pow_half_power(double r2, double power) {
    if (power == 12.0 || power == 6.0) return pow36(r2, power == 12.0);
    const double h = power / 2.0;
    if (h >= 1.5 && h <= 15.5 && h - 0.5 == rint(h - 0.5) && r2 > 0x1p-30 && r2 < 0x1p30) {
        double lo;
        const double hi = half_pow_fast(r2, (int)(h - 0.5), lo);
        if (rounding_settled(hi, lo, 0x1p-92)) return hi + lo;
    }
    return cr_pow(r2, h);
}
ECMC_FN double root_two_over_power(double x, double power) {
    if (power == 12.0 || power == 6.0) return root36(x, power == 12.0);
    const double y = 2.0 / power;
    if (power >= 3.0 && power <= 31.0 && power == rint(power) && ((int)power & 1) && x > 0x1p-300 && x < 0x1p300) {
        double lo;
        const double hi = rational_pow_fast(x, y, 2, (int)power, lo);
        if (rounding_settled(hi, lo, 0x1p-88)) return hi + lo;
    }
    return cr_pow(x, y);
}

template <int D>
ECMC_FN double ip_potential(const double *s, double cp, double power, double prefactor) {
    return cp * prefactor / pow_half_power(norm_sq<D>(s), power);
}

// InversePowerPotential.standard_velocity_derivative (:71-94) times the speed
// (potential/abstracts.py:80-112).
template <int D>
ECMC_FN double ip_derivative(int dir, double speed, const double *s, double c1, double c2, double power,
                             double prefactor) {
    double nrm = cr_pow(norm_sq<D>(s), 0.5);
    double d = power * s[dir] / cr_pow(nrm, power + 2) * prefactor * c1 * c2;
    return d * speed;
}

// InversePowerPotential.standard_velocity_displacement (:96-179) divided by the speed
// (potential/abstracts.py:212-243).
template <int D>
ECMC_FN double ip_displacement(int dir, double speed, const double *sep, double c1, double c2, double power,
                               double prefactor, double dU) {
    double s[D];
#pragma unroll
    for (int i = 0; i < D; i++) s[i] = sep[i];
    double cp = c1 * c2;
    double pp = prefactor * cp;
    if (pp > 0) {
        if (s[dir] <= 0.0) return INFINITY;
        double smax[D];
#pragma unroll
        for (int i = 0; i < D; i++) smax[i] = i == dir ? 0.0 : s[i];
        double maxp = ip_potential<D>(smax, cp, power, prefactor);
        double cur = ip_potential<D>(s, cp, power, prefactor);
        if (dU < maxp - cur) {
            double nsq = root_two_over_power(cp * prefactor / (cur + dU), power);
            double a = nsq - others_sq<D>(s, dir);
            if (a < 0.0) return NAN;
            return (s[dir] - sqrt(a)) / speed;
        }
        return INFINITY;
    }
    double disp = 0.0;
    if (s[dir] > 0.0) {
        disp += s[dir];
        s[dir] = 0.0;
    }
    double cur = ip_potential<D>(s, cp, power, prefactor);
    if (cur + dU >= 0.0) return INFINITY;
    double nsq = root_two_over_power(cp * prefactor / (cur + dU), power);
    double a = nsq - others_sq<D>(s, dir);
    if (a < 0.0) return NAN;
    disp += s[dir] + sqrt(a);
    return disp / speed;
}

// Derivative of the truncated-and-shifted inverse power: the inverse power's inside the
// cutoff sphere (|s| < rc), zero outside.
template <int D>
ECMC_FN double ipt_derivative(int dir, double speed, const double *s, double c1, double c2, double power,
                              double prefactor, double cutoff) {
    if (!(norm_sq<D>(s) < cutoff * cutoff)) return 0.0;
    return ip_derivative<D>(dir, speed, s, c1, c2, power, prefactor);
}

// Truncated-and-shifted inverse power U(r) - U(rc) for r < rc (no reference counterpart):
// the same inversion with the floor U(rc) in place of zero; a target whose line of motion
// stays outside the cutoff sphere never produces an event.
template <int D>
ECMC_FN double ipt_displacement(int dir, double speed, const double *sep, double c1, double c2, double power,
                                double prefactor, double cutoff, double dU) {
    double s[D];
#pragma unroll
    for (int i = 0; i < D; i++) s[i] = sep[i];
    double cp = c1 * c2;
    double pp = prefactor * cp;
    double rc2 = cutoff * cutoff;
    double perp2 = others_sq<D>(s, dir);
    if (perp2 >= rc2) return INFINITY;
    double urc = cp * prefactor / pow_half_power(rc2, power);
    if (pp > 0) {
        if (s[dir] <= 0.0) return INFINITY;
        double s2 = norm_sq<D>(s);
        double cur = s2 < rc2 ? ip_potential<D>(s, cp, power, prefactor) : urc;
        double smax[D];
#pragma unroll
        for (int i = 0; i < D; i++) smax[i] = i == dir ? 0.0 : s[i];
        double maxp = ip_potential<D>(smax, cp, power, prefactor);
        if (dU < maxp - cur) {
            double nsq = root_two_over_power(cp * prefactor / (cur + dU), power);
            double a = nsq - perp2;
            if (a < 0.0) return NAN;
            return (s[dir] - sqrt(a)) / speed;
        }
        return INFINITY;
    }
    double disp = 0.0;
    if (s[dir] > 0.0) {
        disp += s[dir];
        s[dir] = 0.0;
    }
    double s2 = norm_sq<D>(s);
    if (s2 >= rc2) return INFINITY;
    double cur = ip_potential<D>(s, cp, power, prefactor);
    if (cur + dU >= urc) return INFINITY;
    double nsq = root_two_over_power(cp * prefactor / (cur + dU), power);
    double a = nsq - perp2;
    if (a < 0.0) return NAN;
    disp += s[dir] + sqrt(a);
    return disp / speed;
}

// ---- MexicanHatPotential (potential/abstracts.py:285-581) -------------------------------
// LennardJonesPotential (lennard_jones_potential.py:43-135) and DisplacedEvenPowerPotential
// (displaced_even_power_potential.py:44-142).  Helpers return false where the Python code
// raises ValueError (math.sqrt of a negative number); the caller follows the reference's
// try/except in _displacement_behind_outside_sphere.
struct MexicanHat {
    int lj;                 // 1: Lennard-Jones, 0: displaced even power
    double prefactor, sigma, r0, r0sq, k6, k12, power, inverse_power;
};
ECMC_FN MexicanHat mh_lennard_jones(double prefactor, double sigma) {
    MexicanHat h;
    h.lj = 1;
    h.prefactor = prefactor;
    h.sigma = sigma;
    h.r0 = sigma * cr_pow(2.0, 1.0 / 6.0);
    h.r0sq = h.r0 * h.r0;
    h.k6 = -prefactor * cr_pow(sigma, 6.0);
    h.k12 = prefactor * cr_pow(sigma, 12.0);
    h.power = 0; h.inverse_power = 0;
    return h;
}
ECMC_FN MexicanHat mh_displaced_even_power(double r0, double power, double prefactor) {
    MexicanHat h;
    h.lj = 0;
    h.prefactor = prefactor;
    h.sigma = 0;
    h.r0 = r0;
    h.r0sq = r0 * r0;
    h.k6 = 0; h.k12 = 0;
    h.power = power;
    h.inverse_power = 1.0 / power;
    return h;
}
template <int D>
ECMC_FN double mh_potential(const MexicanHat &h, const double *s) {
    if (h.lj) return ip_potential<D>(s, 1.0, 6.0, h.k6) + ip_potential<D>(s, 1.0, 12.0, h.k12);
    double d = cr_pow(norm_sq<D>(s), 0.5) - h.r0;
    return h.prefactor * cr_pow(d, h.power);
}
ECMC_FN double mh_invert_inside(const MexicanHat &h, double u) {
    if (h.lj) {
        double so6 = (1 + cr_pow(1 + 4 * u / h.prefactor, 0.5)) / 2;
        return h.sigma / cr_pow(so6, 1.0 / 6.0);
    }
    return h.r0 - cr_pow(u / h.prefactor, h.inverse_power);
}
ECMC_FN double mh_invert_outside(const MexicanHat &h, double u) {
    if (h.lj) {
        if (u >= 0.0) return INFINITY;
        double so6 = (1 - cr_pow(1 + 4 * u / h.prefactor, 0.5)) / 2;
        return h.sigma / cr_pow(so6, 1.0 / 6.0);
    }
    return h.r0 + cr_pow(u / h.prefactor, h.inverse_power);
}
template <int D>
ECMC_FN bool mh_disp_pos(const double *s, double nsq, int dir, double &out) {
    double a = nsq - others_sq<D>(s, dir);
    if (a < 0.0) return false;
    out = s[dir] - sqrt(a);
    return true;
}
template <int D>
ECMC_FN bool mh_disp_neg(const double *s, double nsq, int dir, double &out) {
    double a = nsq - others_sq<D>(s, dir);
    if (a < 0.0) return false;
    out = s[dir] + sqrt(a);
    return true;
}
template <int D>
ECMC_FN bool mh_front_outside(const MexicanHat &h, int dir, double cur, double dU, double *s, double &out) {
    double nn = mh_invert_outside(h, cur + dU);
    return mh_disp_neg<D>(s, nn * nn, dir, out);
}
template <int D>
ECMC_FN bool mh_front_inside(const MexicanHat &h, int dir, double dU, double *s, double &out) {
    double disp, more;
    if (!mh_disp_neg<D>(s, h.r0sq, dir, disp)) return false;
    s[dir] -= disp;
    double cur = mh_potential<D>(h, s);
    if (!mh_front_outside<D>(h, dir, cur, dU, s, more)) return false;
    out = disp + more;
    return true;
}
template <int D>
ECMC_FN bool mh_behind_inside(const MexicanHat &h, int dir, double cur, double dU, double *s, double &out) {
    double smax[D];
#pragma unroll
    for (int i = 0; i < D; i++) smax[i] = i == dir ? 0.0 : s[i];
    double diff = mh_potential<D>(h, smax) - cur;
    if (dU < diff) {
        double nn = mh_invert_inside(h, cur + dU);
        return mh_disp_pos<D>(s, nn * nn, dir, out);
    }
    double disp = s[dir], more;
    s[dir] = 0.0;
    dU -= diff;
    if (!mh_front_inside<D>(h, dir, dU, s, more)) return false;
    out = disp + more;
    return true;
}
template <int D>
ECMC_FN bool mh_behind_outside(const MexicanHat &h, int dir, double dU, double *s, double &out) {
    double disp, more;
    if (mh_disp_pos<D>(s, h.r0sq, dir, disp)) {
        s[dir] -= disp;
        double cur = mh_potential<D>(h, s);
        if (mh_behind_inside<D>(h, dir, cur, dU, s, more)) {
            out = disp + more;
            return true;
        }
    }
    disp = s[dir];
    s[dir] = 0.0;
    double cur = mh_potential<D>(h, s);
    if (!mh_front_outside<D>(h, dir, cur, dU, s, more)) return false;
    out = disp + more;
    return true;
}
template <int D>
ECMC_FN double mh_displacement(const MexicanHat &h, int dir, double speed, const double *sep, double dU) {
    double s[D], out = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++) s[i] = sep[i];
    double nrm = cr_pow(norm_sq<D>(s), 0.5);
    bool ok;
    if (nrm >= h.r0) {
        if (s[dir] <= 0.0) {
            double cur = mh_potential<D>(h, s);
            ok = mh_front_outside<D>(h, dir, cur, dU, s, out);
        } else {
            ok = mh_behind_outside<D>(h, dir, dU, s, out);
        }
    } else {
        if (s[dir] <= 0.0) {
            ok = mh_front_inside<D>(h, dir, dU, s, out);
        } else {
            double cur = mh_potential<D>(h, s);
            ok = mh_behind_inside<D>(h, dir, cur, dU, s, out);
        }
    }
    if (!ok) return NAN;
    return out / speed;
}
template <int D>
ECMC_FN double mh_derivative(const MexicanHat &h, int dir, double speed, const double *s) {
    if (h.lj) {
        double d = ip_derivative<D>(dir, 1.0, s, 1.0, 1.0, 6.0, h.k6) + ip_derivative<D>(dir, 1.0, s, 1.0, 1.0, 12.0, h.k12);
        return d * speed;
    }
    double nrm = cr_pow(norm_sq<D>(s), 0.5);
    double d = -h.power * h.prefactor * cr_pow(nrm - h.r0, h.power - 1) * s[dir] / nrm;
    return d * speed;
}

// ---- InversePowerCoulombBoundingPotential -----------------------------------------------
// (inverse_power_coulomb_bounding_potential.c:53-139); x is the direction of motion.
ECMC_FN double ipcb_derivative(double pp, double sx, double sy, double sz) {
    return pp * sx / cr_pow(sx * sx + sy * sy + sz * sz, 3.0 / 2.0);
}
ECMC_FN double ipcb_potential(double pp, double sx, double sy, double sz) {
    return pp / sqrt(sx * sx + sy * sy + sz * sz);
}
// Windings of the image sum first (floor/fmod), then the closed-form remainder.
ECMC_FN double ipcb_displacement(double pp, double sx, double sy, double sz, double dU, double L) {
    double half = L / 2.0;
    double cur = ipcb_potential(pp, sx, sy, sz);
    double p0 = ipcb_potential(pp, 0.0, sy, sz);
    double ph = ipcb_potential(pp, half, sy, sz);
    double per = fabs(p0 - ph);
    double disp = floor(dU / per) * L;
    dU = fmod_near(dU, per);
    double nn;
    if (pp > 0.0) {
        if (sx <= 0.0) {
            disp += half + sx;
            sx = half;
            cur = ph;
        } else if (dU >= p0 - cur) {
            dU -= (p0 - cur);
            disp += sx + half;
            sx = half;
            cur = ph;
        }
        nn = pp / (cur + dU);
        disp += (sx - sqrt(nn * nn - (sy * sy + sz * sz)));
    } else {
        if (sx > 0.0) {
            disp += sx;
            sx = 0.0;
            cur = p0;
        } else if (dU >= ph - cur) {
            dU -= (ph - cur);
            disp += sx + L;
            sx = 0.0;
            cur = p0;
        }
        nn = pp / (cur + dU);
        disp += (sx + sqrt(nn * nn - (sy * sy + sz * sz)));
    }
    return disp;
}

// ---- MergedImageCoulombPotential derivative along +x ------------------------------------
// (merged_image_coulomb_potential.c:205-274): real-space images |n| <= position_cutoff in
// the same k, j, i order, then the Fourier sum over the table of build_ewald_table with the
// same cos/sin recurrence, so that the summation rounds as the reference's does.
struct Ewald {
    const double *f;
    int fc, pc;
    double L, aol, aol2, two_aol_rpi, tpol;
};
ECMC_FN double ewald_derivative(const Ewald &e, double sx, double sy, double sz) {
    double der = 0.0;
    const int pc = e.pc, pc2 = pc * pc, fc = e.fc, fc2 = fc * fc, n = fc + 1;
    const double L = e.L;
    for (int k = -pc; k < pc + 1; k++) {
        double vz2 = (sz + k * L) * (sz + k * L);
        int cy = (int)sqrt((double)(pc2 - k * k));
        for (int j = -cy; j < cy + 1; j++) {
            double vy2 = (sy + j * L) * (sy + j * L);
            int cx = (int)sqrt((double)(pc2 - j * j - k * k));
            for (int i = -cx; i < cx + 1; i++) {
                double vx = sx + i * L;
                double v2 = vx * vx + vy2 + vz2;
                double vn = sqrt(v2);
                der += vx * (e.two_aol_rpi * exp(-e.aol2 * v2) + erfc(e.aol * vn) / vn) / v2;
            }
        }
    }
    double dcx = cos(e.tpol * sx), dsx = sin(e.tpol * sx);
    double dcy = cos(e.tpol * sy), dsy = sin(e.tpol * sy);
    double dcz = cos(e.tpol * sz), dsz = sin(e.tpol * sz);
    double cxv = dcx, sxv = dsx, cyv = 1.0, syv = 0.0, czv = 1.0, szv = 0.0, tmp;
    for (int i = 1; i < fc + 1; i++) {
        int cy = (int)sqrt((double)(fc2 - i * i));
        for (int j = 0; j < cy + 1; j++) {
            int cx = (int)sqrt((double)(fc2 - i * i - j * j));
            for (int k = 0; k < cx + 1; k++) {
                der += e.f[(i * n + j) * n + k] * sxv * cyv * czv;
                if (k != cx) {
                    tmp = czv; czv = tmp * dcz - szv * dsz; szv = szv * dcz + tmp * dsz;
                } else if (j != cy) {
                    tmp = cyv; cyv = tmp * dcy - syv * dsy; syv = syv * dcy + tmp * dsy;
                    czv = 1.0; szv = 0.0;
                } else if (i != fc) {
                    tmp = cxv; cxv = tmp * dcx - sxv * dsx; sxv = sxv * dcx + tmp * dsx;
                    cyv = 1.0; syv = 0.0; czv = 1.0; szv = 0.0;
                }
            }
        }
    }
    return der;
}

}  // namespace ecmc
