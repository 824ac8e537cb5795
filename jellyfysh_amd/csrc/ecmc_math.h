// ecmc_math.h -- FP64 arithmetic of the ECMC hot path for gfx950.
//
// * Python float semantics the reference relies on: `x % L` (setting/hypercubic_setting.py:117,172)
//   and divmod(x, 1.0) inside base/time.py:101,129-131.  Both follow CPython's float_rem /
//   float_divmod rules (remainder takes the divisor's sign, quotient snapped to an integer).
// * Correctly rounded log and pow.  The reference evaluates `x ** y` and math.log through
//   the C library; a correctly rounded result is the one an exact evaluation rounds to, so
//   every platform that rounds correctly agrees bit for bit.  They are computed here in
//   double-double arithmetic (about 2^-100 relative error before the final rounding):
//     log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), m r_i - 1 = z exactly in double-double with
//          r_i ~ 1/(1+i/128) from ecmc_tables.h, log1p(z) = 2 atanh(z/(2+z)) by its Taylor
//          series, plus e ln2 and -log(r_i) from the table.
//     exp: t = k ln2/64 + r, |r| <= ln2/128, Taylor series of exp(r) to r^11, times the
//          double-double 2^(j/64) table and 2^(k>>6).
//   pow(x, y) = exp(y log x) with both steps in double-double.
//
// Everything here is compiled with -ffp-contract=off: fused operations appear only where
// fma() is written out (the exact product error of two_prod).
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__) && !defined(ECMC_HOST_BUILD)
#include <hip/hip_runtime.h>
#define ECMC_FN __device__ __forceinline__
#define ECMC_HD __host__ __device__ __forceinline__
#define ECMC_TABLE __constant__
// the rarely taken double-double fallbacks: called, not inlined, so that their registers do
// not add to those of the fast paths around every call site
#define ECMC_SLOW static __device__ __noinline__
#define ECMC_HD_SLOW static __host__ __device__ __noinline__
#else
#define ECMC_FN static inline
#define ECMC_HD static inline
#define ECMC_TABLE
#define ECMC_SLOW static
#define ECMC_HD_SLOW static
#endif

#include "ecmc_tables.h"

namespace ecmc {

// ---- Python float semantics ------------------------------------------------------------
// fmod beyond one period (a reduction loop): called, not inlined, where the one-period paths
// below leave it
ECMC_HD_SLOW double fmod_call(double x, double y) { return fmod(x, y); }

ECMC_HD double py_mod(double x, double y) {
    if (y > 0.0 && fabs(x) < 2.0 * y) {
        // within one period either way fmod's exact result is x, x - y or x + y (Sterbenz), and a
        // negative remainder takes the divisor's sign by one rounded addition, as below
        double m = x >= y ? x - y : (x <= -y ? x + y : x);
        if (m < 0.0) m += y;
        return m == 0.0 ? 0.0 : m;
    }
    double m = fmod_call(x, y);
    if (m != 0.0) {
        if ((y < 0) != (m < 0)) m += y;
    } else {
        m = copysign(0.0, y);
    }
    return m;
}

// fmod(x, y), with the one-period cases taken without fmod's reduction loop: x for
// 0 <= x < y, the exact x - y for y <= x < 2y (Sterbenz).
ECMC_HD double fmod_near(double x, double y) {
    if (x >= 0.0 && y > 0.0 && x < 2.0 * y) return x < y ? x : x - y;
    return fmod_call(x, y);
}

ECMC_HD void py_divmod(double vx, double wx, double &q, double &r) {
    if (wx == 1.0 && vx >= 0.0 && vx < 0x1p52) {
        // division by one of a non-negative double: the floor and the exact remainder (+0 for an
        // integer, a -0 quotient for -0 as copysign gives below)
        const double f = floor(vx);
        q = f;
        r = vx - f;
        return;
    }
    double mod = fmod_call(vx, wx);
    double div = (vx - mod) / wx;
    if (mod != 0.0) {
        if ((wx < 0) != (mod < 0)) { mod += wx; div -= 1.0; }
    } else {
        mod = copysign(0.0, wx);
    }
    double fl;
    if (div != 0.0) {
        fl = floor(div);
        if (div - fl > 0.5) fl += 1.0;
    } else {
        fl = copysign(0.0, vx / wx);
    }
    q = fl;
    r = mod;
}

// Event time as (quotient, remainder) of a division by one (base/time.py:87-182).
struct Time {
    double q, r;
};
ECMC_HD Time time_add(Time t, double dt) {
    Time o;
    if (!isinf(dt)) {
        double aq, nr;
        py_divmod(t.r + dt, 1.0, aq, nr);
        o.q = t.q + aq;
        o.r = nr;
    } else {
        o.q = dt;
        o.r = dt;
    }
    return o;
}
ECMC_HD double time_sub(Time a, Time b) { return a.q - b.q + a.r - b.r; }
ECMC_HD bool time_lt(Time a, Time b) { return a.q < b.q || (a.q == b.q && a.r < b.r); }
ECMC_HD bool time_eq(Time a, Time b) { return a.q == b.q && a.r == b.r; }
ECMC_HD Time time_from_float(double t) {
    Time o;
    if (isinf(t)) { o.q = t; o.r = t; } else { py_divmod(t, 1.0, o.q, o.r); }
    return o;
}

// ---- double-double ----------------------------------------------------------------------
struct dd {
    double h, l;
};
ECMC_FN dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
ECMC_FN dd quick_two_sum(double a, double b) {
    double s = a + b;
    double e = b - (s - a);
    return {s, e};
}
ECMC_FN dd two_prod(double a, double b) {
    double p = a * b;
    double e = fma(a, b, -p);
    return {p, e};
}
ECMC_FN dd dd_add(dd a, dd b) {
    dd s = two_sum(a.h, b.h);
    dd t = two_sum(a.l, b.l);
    s.l += t.h;
    s = quick_two_sum(s.h, s.l);
    s.l += t.l;
    return quick_two_sum(s.h, s.l);
}
ECMC_FN dd dd_neg(dd a) { return {-a.h, -a.l}; }
ECMC_FN dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }
ECMC_FN dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.h, b.h);
    p.l += a.h * b.l + a.l * b.h;
    return quick_two_sum(p.h, p.l);
}
ECMC_FN dd dd_mul_d(dd a, double b) {
    dd p = two_prod(a.h, b);
    p.l += a.l * b;
    return quick_two_sum(p.h, p.l);
}
ECMC_FN dd dd_div(dd a, dd b) {
    double q1 = a.h / b.h;
    dd r = dd_sub(a, dd_mul_d(b, q1));
    double q2 = r.h / b.h;
    r = dd_sub(r, dd_mul_d(b, q2));
    double q3 = r.h / b.h;
    dd q = quick_two_sum(q1, q2);
    return dd_add(q, dd{q3, 0.0});
}

// Natural logarithm of a positive, finite, normal x in double-double.
ECMC_FN dd dd_log(double x) {
    int e;
    double m = frexp(x, &e);          // m in [0.5, 1)
    m = m * 2.0;
    e -= 1;                            // m in [1, 2)
    if (m > 0x1.6a09e667f3bcdp+0) {   // sqrt(2)
        m = m * 0.5;
        e += 1;
    }
    int i = (int)rint((m - 1.0) * 128.0);
    int ti = i - ECMC_LOG_TMIN;
    double r = ecmc_log_r[ti];
    dd p = two_prod(m, r);
    dd z = quick_two_sum(p.h - 1.0, p.l);    // p.h - 1 is exact
    dd w = dd_div(z, dd_add(dd{2.0, 0.0}, z));
    dd s = dd_mul(w, w);
    double sh = s.h;
    double tail = (((1.0 / 15.0) * sh + 1.0 / 13.0) * sh + 1.0 / 11.0) * sh + 1.0 / 9.0;
    dd acc = {tail, 0.0};
    acc = dd_add(dd{ECMC_INV7_H, ECMC_INV7_L}, dd_mul(s, acc));
    acc = dd_add(dd{ECMC_INV5_H, ECMC_INV5_L}, dd_mul(s, acc));
    acc = dd_add(dd{ECMC_INV3_H, ECMC_INV3_L}, dd_mul(s, acc));
    acc = dd_add(dd{1.0, 0.0}, dd_mul(s, acc));
    dd l1p = dd_mul(dd_mul_d(w, 2.0), acc);
    double ed = (double)e;
    dd el = two_prod(ed, ECMC_LN2_H);
    el = dd_add(el, dd{ed * ECMC_LN2_L, 0.0});
    dd t = {ecmc_log_th[ti], ecmc_log_tl[ti]};
    return dd_add(dd_add(el, t), l1p);
}

// exp of a double-double argument with |t| < 700.
ECMC_FN dd dd_exp(dd t) {
    double kf = rint(t.h * ECMC_INV_LN2_64);
    int k = (int)kf;
    dd r = two_sum(t.h - kf * ECMC_LN2_64_HI, t.l);
    r = dd_sub(r, two_prod(kf, ECMC_LN2_64_MID));
    r = dd_sub(r, dd{kf * ECMC_LN2_64_LO, 0.0});
    double rh = r.h;
    double tail = ((((ECMC_INVFACT12_H * rh + ECMC_INVFACT11_H) * rh + ECMC_INVFACT10_H) * rh + ECMC_INVFACT9_H) * rh
                   + ECMC_INVFACT8_H) * rh + ECMC_INVFACT7_H;
    tail = tail * rh + ECMC_INVFACT6_H;
    dd acc = {tail, 0.0};
    acc = dd_add(dd{ECMC_INVFACT5_H, ECMC_INVFACT5_L}, dd_mul(r, acc));
    acc = dd_add(dd{ECMC_INVFACT4_H, ECMC_INVFACT4_L}, dd_mul(r, acc));
    acc = dd_add(dd{ECMC_INVFACT3_H, ECMC_INVFACT3_L}, dd_mul(r, acc));
    acc = dd_add(dd{0.5, 0.0}, dd_mul(r, acc));
    acc = dd_add(dd{1.0, 0.0}, dd_mul(r, acc));
    acc = dd_add(dd{1.0, 0.0}, dd_mul(r, acc));
    int j = k & 63;
    int sc = (k - j) / 64;
    dd res = dd_mul(acc, dd{ecmc_exp2_h[j], ecmc_exp2_l[j]});
    res.h = ldexp(res.h, sc);
    res.l = ldexp(res.l, sc);
    return res;
}

// True when every value within hi + lo +- rel * |hi| rounds to the same double as hi + lo:
// the correctly rounded result of a quantity known to that accuracy is then hi + lo.
ECMC_FN bool rounding_settled(double hi, double lo, double rel) {
    const double e = rel * fabs(hi);
    return hi + (lo - e) == hi + (lo + e);
}

// log x as hi + lo with relative error below 2^-64 (2^-68.5 the largest seen against quad
// precision), for a positive normal x (the fast path of cr_log).  Same reduction as dd_log, x = 2^e m, z = m r_i - 1 = zh + zl exactly, |zh| < 2^-7;
// log1p(z) = z - z^2/2 + sum_{k>=3} (-1)^(k+1) z^k / k, whose first two terms are summed
// exactly and whose tail zh^3 (1/3 - zh/4 + ... - zh^7/10) (truncation below zh^11/11) and the
// first-order zl terms are taken in double.  Relative to |log x| >= 2^-8 (i != 0 or e != 0),
// or to log1p(zh) when i = e = 0 (zl = 0 there), the double parts contribute below 2^-67.
ECMC_FN double log_fast(double x, double &lo) {
    int e;
    double m = frexp(x, &e);
    m = m * 2.0;
    e -= 1;
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    const int ti = (int)rint((m - 1.0) * 128.0) - ECMC_LOG_TMIN;
    const dd p = two_prod(m, ecmc_log_r[ti]);
    const double zh = p.h - 1.0, zl = p.l;      // p.h - 1 is exact
    const dd z2 = two_prod(zh, zh);
    double q = -1.0 / 10.0;
    q = fma(q, zh, 1.0 / 9.0);
    q = fma(q, zh, -1.0 / 8.0);
    q = fma(q, zh, 1.0 / 7.0);
    q = fma(q, zh, -1.0 / 6.0);
    q = fma(q, zh, 1.0 / 5.0);
    q = fma(q, zh, -1.0 / 4.0);
    q = fma(q, zh, 1.0 / 3.0);
    const double tail = (zh * z2.h) * q + zl * (z2.h * (1.0 - zh) - zh);
    const double ed = (double)e;
    const dd el = two_prod(ed, ECMC_LN2_H);
    const dd s1 = two_sum(el.h, ecmc_log_th[ti]);
    const dd s2 = two_sum(s1.h, zh);
    const dd s3 = two_sum(s2.h, -0.5 * z2.h);
    const double l = ((s1.l + s2.l) + s3.l) +
                     (((el.l + ed * ECMC_LN2_L) + ecmc_log_tl[ti]) + ((zl - 0.5 * z2.l) + tail));
    const dd r = quick_two_sum(s3.h, l);
    lo = r.l;
    return r.h;
}

ECMC_SLOW double cr_log_slow(double x) { return dd_log(x).h; }
// log of zero, negative, infinite, NaN and subnormal arguments
ECMC_SLOW double log_special(double x) { return log(x); }

// Correctly rounded natural logarithm: the fast path when its rounding is settled (all but
// about one call in 500), else the double-double path.
ECMC_FN double cr_log(double x) {
    if (!(x > 0.0) || isinf(x) || x < 0x1p-1022) return log_special(x);
    if (x == 1.0) return 0.0;
    double lo;
    const double hi = log_fast(x, lo);
    if (rounding_settled(hi, lo, 0x1p-64)) return hi + lo;
    return cr_log_slow(x);
}

// x ** y for y the double nearest 1/k, k = 3 or 6 (2/p of the inverse-power inversion for
// p = 6, 12), as hi + lo with relative error below 2^-92.  t0 = x^(1/k) from cbrt (of sqrt x
// for k = 6) to about 2^-51; one Newton step t1 = t0 - t0 (t0^k - x) / (k x) with t0^k - x in
// double-double leaves (k - 1)/2 (t0 error)^2; then x^y = t1 x^delta = t1 (1 + delta log x +
// O(2^-104)) with delta = y - 1/k = (k y - 1)/k, |delta log x| < 2^-48 for x < 2^1000, log x
// from the fast path above.  k = 3 and 6 share one instruction stream (selects, no branches),
// so lanes of a wavefront evaluating different powers do not diverge.
ECMC_FN double root_pow_fast(double x, bool six, double &lo) {
    const double k = six ? 6.0 : 3.0;
    const double y = six ? 1.0 / 6.0 : 1.0 / 3.0;
    const double t0 = cbrt(six ? sqrt(x) : x);
    const dd t2 = two_prod(t0, t0);
    const dd t3 = dd_mul_d(t2, t0);
    const dd t6 = dd_mul(t3, t3);
    const dd tk = six ? t6 : t3;
    const dd dlt = dd_sub(tk, dd{x, 0.0});
    const double corr = t0 * ((dlt.h + dlt.l) / (k * x));
    const dd t1 = two_sum(t0, -corr);
    const double delta = fma(y, k, -1.0) / k;
    double llo;
    const double c = delta * log_fast(x, llo);
    const dd r = quick_two_sum(t1.h, fma(t1.h, c, t1.l));
    lo = r.l;
    return r.h;
}

// x ** n for an integer 3 <= n <= 32 by binary powering in double-double (each product
// carries about 2^-104 relative error), rounded once.
ECMC_FN double dd_pow_int(double x, int n) {
    dd base = {x, 0.0};
    dd acc = {1.0, 0.0};
    bool first = true;
    while (n > 0) {
        if (n & 1) {
            acc = first ? base : dd_mul(acc, base);
            first = false;
        }
        n >>= 1;
        if (n) base = dd_mul(base, base);
    }
    return acc.h;
}

// x ** n as a double-double by binary powering (each product about 2^-104 relative), n >= 1.
ECMC_FN dd dd_pow_int_dd(dd base, int n) {
    dd acc = {1.0, 0.0};
    bool first = true;
    while (n > 0) {
        if (n & 1) {
            acc = first ? base : dd_mul(acc, base);
            first = false;
        }
        n >>= 1;
        if (n) base = dd_mul(base, base);
    }
    return acc;
}

// x ** (n + 1/2) for 1 <= n <= 15 as hi + lo, relative error below 2^-98: x^n by binary powering
// in double-double, sqrt(x) = s + (x - s^2) / (2 s) with the residual exact by fma.
ECMC_FN double half_pow_fast(double x, int n, double &lo) {
    const double s = sqrt(x);
    const dd sq = quick_two_sum(s, fma(-s, s, x) / (2.0 * s));
    const dd r = dd_mul(dd_pow_int_dd(dd{x, 0.0}, n), sq);
    lo = r.l;
    return r.h;
}

// x ** y for y the double nearest num / den (num = 1, 2; 2 <= den <= 32) as hi + lo, relative
// error below 2^-92: t0 = pow(x, y) (within a few ulps), one Newton step on t^den = x^num with
// both powers in double-double, t1 = t0 - t0 (t0^den - x^num) / (den x^num), which leaves
// (den - 1)/2 (t0 error)^2; then x^y = t1 (1 + delta log x) with delta = y - num/den =
// (den y - num)/den, |delta log x| < 2^-45 for |log2 x| < 300.
ECMC_FN double rational_pow_fast(double x, double y, int num, int den, double &lo) {
    const double t0 = pow(x, y);
    const dd tk = dd_pow_int_dd(dd{t0, 0.0}, den);
    const dd xn = num == 1 ? dd{x, 0.0} : two_prod(x, x);
    const dd dlt = dd_sub(tk, xn);
    const double corr = t0 * ((dlt.h + dlt.l) / (den * xn.h));
    const dd t1 = two_sum(t0, -corr);
    const double delta = fma(y, (double)den, -(double)num) / den;
    double llo;
    const double c = delta * log_fast(x, llo);
    const dd r = quick_two_sum(t1.h, fma(t1.h, c, t1.l));
    lo = r.l;
    return r.h;
}

// Correctly rounded x ** y for the arguments the potentials use (x > 0): the cases that need
// more than one operation.  Fast paths with a rounding test first (integer powers, half-integer
// powers r^(p/2) of odd p, and 2/p for the inversion of an integer power p), then the
// double-double logarithm and exponential.
ECMC_SLOW double cr_pow_slow(double x, double y) {
    if (!(x > 0.0) || isinf(x) || !isfinite(y) || x < 0x1p-1022 || x == 1.0) return pow(x, y);
    if (y >= 3.0 && y <= 32.0 && y == rint(y) && x > 0x1p-30 && x < 0x1p30) return dd_pow_int(x, (int)y);
    if ((y == 1.0 / 3.0 || y == 1.0 / 6.0) && x > 0x1p-900 && x < 0x1p900) {
        double lo;
        const double hi = root_pow_fast(x, y == 1.0 / 6.0, lo);
        if (rounding_settled(hi, lo, 0x1p-92)) return hi + lo;
    }
    if (y >= 1.5 && y <= 15.5 && y - 0.5 == rint(y - 0.5) && x > 0x1p-30 && x < 0x1p30) {
        double lo;
        const double hi = half_pow_fast(x, (int)(y - 0.5), lo);
        if (rounding_settled(hi, lo, 0x1p-92)) return hi + lo;
    }
    if (y > 0.0 && y < 1.0 && x > 0x1p-300 && x < 0x1p300) {
        // y the double nearest 2/p (p odd, reduced to 2/p) or 1/q (p = 2q)
        const double pd = rint(2.0 / y);
        if (pd >= 2.0 && pd <= 32.0 && 2.0 / pd == y) {
            const int pi = (int)pd;
            const int num = (pi & 1) ? 2 : 1, den = (pi & 1) ? pi : pi / 2;
            if (den >= 2) {
                double lo;
                const double hi = rational_pow_fast(x, y, num, den, lo);
                if (rounding_settled(hi, lo, 0x1p-88)) return hi + lo;
            }
        }
    }
    dd l = dd_log(x);
    dd t = dd_mul_d(l, y);
    if (fabs(t.h) > 700.0) return pow(x, y);
    dd e = dd_exp(t);
    return e.h;
}

ECMC_FN double cr_pow(double x, double y) {
    if (y == 2.0) return x * x;
    if (y == 0.5 && x >= 0.0) return x == 0.0 ? 0.0 : sqrt(x);
    if (y == 1.0) return x;
    if (y == 0.0) return 1.0;
    return cr_pow_slow(x, y);
}

// x ** 3 or x ** 6 (six) in double-double, rounded once: x^2 exact, x^3 = x^2 x and
// x^6 = (x^3)^2 each about 2^-104 relative, for 2^-30 < x < 2^30 (else cr_pow).
ECMC_FN double pow36(double x, bool six) {
    if (!(x > 0x1p-30 && x < 0x1p30)) return cr_pow_slow(x, six ? 6.0 : 3.0);
    const dd x2 = two_prod(x, x);
    const dd x3 = dd_mul_d(x2, x);
    const dd x6 = dd_mul(x3, x3);
    return six ? x6.h : x3.h;
}

// Correctly rounded x ** (1/3) or x ** (1/6) for the double exponents 1.0/3.0 and 1.0/6.0.
ECMC_FN double root36(double x, bool six) {
    if (x > 0x1p-900 && x < 0x1p900) {
        double lo;
        const double hi = root_pow_fast(x, six, lo);
        if (rounding_settled(hi, lo, 0x1p-92)) return hi + lo;
    }
    return cr_pow_slow(x, six ? 1.0 / 6.0 : 1.0 / 3.0);
}

}  // namespace ecmc
