/*
 * ecmc_oracle_math.c -- RNG, Python float semantics and potentials of the oracle.
 * TEST INFRASTRUCTURE ONLY (see ecmc_oracle.h).  Each function names the reference
 * file:line it restates.  Built with -ffp-contract=off so that every expression rounds
 * exactly as CPython evaluates it.
 */
#include "ecmc_oracle.h"
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

/* ===== MT19937 with CPython's seeding and transforms ==================================
 * Matsumoto & Nishimura (1998), init_by_array variant; CPython seeds an int by splitting
 * abs(seed) into 32-bit little-endian words (Modules/_randommodule.c random_seed).
 * random() = (a*2^26+b)/2^53 with a = w>>5, b = w>>6 (genrand_res53). */
#define MT_N 624
#define MT_M 397
static void mt_init_genrand(ora_mt *r, uint32_t s) {
    r->mt[0] = s;
    for (int i = 1; i < MT_N; i++)
        r->mt[i] = 1812433253U * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
    r->mti = MT_N;
}

void ora_mt_seed(ora_mt *r, uint64_t seed) {
    uint32_t key[2];
    int klen;
    key[0] = (uint32_t)(seed & 0xffffffffu);
    key[1] = (uint32_t)(seed >> 32);
    klen = key[1] ? 2 : 1;
    mt_init_genrand(r, 19650218U);
    int i = 1, j = 0;
    int k = MT_N > klen ? MT_N : klen;
    for (; k; k--) {
        r->mt[i] = (r->mt[i] ^ ((r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) * 1664525U)) + key[j] + (uint32_t)j;
        i++; j++;
        if (i >= MT_N) { r->mt[0] = r->mt[MT_N - 1]; i = 1; }
        if (j >= klen) j = 0;
    }
    for (k = MT_N - 1; k; k--) {
        r->mt[i] = (r->mt[i] ^ ((r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) * 1566083941U)) - (uint32_t)i;
        i++;
        if (i >= MT_N) { r->mt[0] = r->mt[MT_N - 1]; i = 1; }
    }
    r->mt[0] = 0x80000000U;
    r->mti = MT_N;
    r->words = 0;
}

uint32_t ora_mt_u32(ora_mt *r) {
    static const uint32_t mag01[2] = {0x0U, 0x9908b0dfU};
    uint32_t y;
    if (r->mti >= MT_N) {
        int kk;
        for (kk = 0; kk < MT_N - MT_M; kk++) {
            y = (r->mt[kk] & 0x80000000U) | (r->mt[kk + 1] & 0x7fffffffU);
            r->mt[kk] = r->mt[kk + MT_M] ^ (y >> 1) ^ mag01[y & 1U];
        }
        for (; kk < MT_N - 1; kk++) {
            y = (r->mt[kk] & 0x80000000U) | (r->mt[kk + 1] & 0x7fffffffU);
            r->mt[kk] = r->mt[kk + (MT_M - MT_N)] ^ (y >> 1) ^ mag01[y & 1U];
        }
        y = (r->mt[MT_N - 1] & 0x80000000U) | (r->mt[0] & 0x7fffffffU);
        r->mt[MT_N - 1] = r->mt[MT_M - 1] ^ (y >> 1) ^ mag01[y & 1U];
        r->mti = 0;
    }
    y = r->mt[r->mti++];
    r->words++;
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680U;
    y ^= (y << 15) & 0xefc60000U;
    y ^= (y >> 18);
    return y;
}

double ora_random(ora_mt *r) {
    uint32_t a = ora_mt_u32(r) >> 5, b = ora_mt_u32(r) >> 6;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

/* random.py:639-654 (3.10): -log(1.0 - random()) / lambd */
double ora_expovariate(ora_mt *r, double lambd, int math_mode) {
    double u = ora_random(r);
    return -ora_log(1.0 - u, math_mode) / lambd;
}

/* random.py:546-548 */
double ora_uniform(ora_mt *r, double a, double b) { return a + (b - a) * ora_random(r); }

/* random.py:366-370 -> randrange -> _randbelow_with_getrandbits (k = n.bit_length(),
 * getrandbits(k) = w >> (32-k) for k <= 32, rejection while r >= n). */
int64_t ora_randint(ora_mt *r, int64_t a, int64_t b) {
    uint64_t n = (uint64_t)(b - a + 1);
    if (n == 0) return a;
    int k = 0;
    while ((n >> k) != 0) k++;
    uint64_t v;
    if (k <= 32) {
        do { v = ora_mt_u32(r) >> (32 - k); } while (v >= n);
    } else {
        do {
            uint64_t lo = ora_mt_u32(r);
            uint64_t hi = ora_mt_u32(r) >> (64 - k);
            v = lo | (hi << 32);
        } while (v >= n);
    }
    return a + (int64_t)v;
}

/* ===== Python float semantics (Objects/floatobject.c float_rem / float_divmod) ===== */
double ora_py_mod(double x, double y) {
    double mod = fmod(x, y);
    if (mod != 0.0) {
        if ((y < 0) != (mod < 0)) mod += y;
    } else {
        mod = copysign(0.0, y);
    }
    return mod;
}

void ora_py_divmod(double vx, double wx, double *q, double *r) {
    double mod = fmod(vx, wx);
    double div = (vx - mod) / wx;
    double floordiv;
    if (mod != 0.0) {
        if ((wx < 0) != (mod < 0)) { mod += wx; div -= 1.0; }
    } else {
        mod = copysign(0.0, wx);
    }
    if (div != 0.0) {
        floordiv = floor(div);
        if (div - floordiv > 0.5) floordiv += 1.0;
    } else {
        floordiv = copysign(0.0, vx / wx);
    }
    *q = floordiv;
    *r = mod;
}

/* base/time.py:115-133 */
void ora_time_add(double q, double r, double dt, double *oq, double *or_) {
    if (!isinf(dt)) {
        double aq, nr;
        ora_py_divmod(r + dt, 1.0, &aq, &nr);
        *oq = q + aq;
        *or_ = nr;
    } else {
        *oq = dt;
        *or_ = dt;
    }
}

/* base/time.py:135-149 */
double ora_time_sub(double q1, double r1, double q2, double r2) { return q1 - q2 + r1 - r2; }


/* ===== math switch ===================================================================== */
double ora_pow(double x, double y, int math_mode) {
    if (math_mode == ORA_MATH_CR) {
        if (y == 2.0) return x * x;               /* glibc agrees exactly for y == 2 */
        if (y == 0.5) return sqrt(x);             /* correctly rounded square root */
        if (x == 0.0 || isinf(x) || isnan(x) || isnan(y)) return pow(x, y);
        return (double)powq((__float128)x, (__float128)y);
    }
    return pow(x, y);
}

double ora_log(double x, int math_mode) {
    if (math_mode == ORA_MATH_CR && x > 0.0 && !isinf(x)) return (double)logq((__float128)x);
    return log(x);
}

/* Array forms of ora_log / ora_pow, for checks over millions of arguments. */
void ora_log_n(int64_t n, const double *x, double *out, int math_mode) {
    for (int64_t i = 0; i < n; i++) out[i] = ora_log(x[i], math_mode);
}
void ora_pow_n(int64_t n, const double *x, const double *y, double *out, int math_mode) {
    for (int64_t i = 0; i < n; i++) out[i] = ora_pow(x[i], y[i], math_mode);
}

/* ===== vectors (base/vectors.py) ======================================================== */
static double v_norm_sq(int dim, const double *s) {
    double acc = 0.0;             /* sum() starts from int 0 */
    for (int i = 0; i < dim; i++) acc = acc + s[i] * s[i];
    return acc;
}
static double v_norm(int dim, const double *s, int mm) { return ora_pow(v_norm_sq(dim, s), 0.5, mm); }

/* SeparationOutputHandler.write for point masses (separation_output_handler.py:81-95): for
 * i < j in order, norm(separation_vector(x_i, x_j)) with the entries corrected as in
 * hypercubic_setting.py:156-172 and the norm of vectors.py:29-43. */
void ora_separation_norms(int dim, int64_t n, const double *pos, double L, double *out, int mm) {
    const double half = L / 2.0;
    int64_t k = 0;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) {
            double s[3];
            for (int d = 0; d < dim; d++) s[d] = ora_py_mod((pos[dim * j + d] - pos[dim * i + d]) + half, L) - half;
            out[k++] = v_norm(dim, s, mm);
        }
}
static double v_others_sq(int dim, const double *s, int dir, int mm) {
    /* sum([value ** 2 for index, value in enumerate(old_vector) if index != direction]) */
    double acc = 0.0;
    for (int i = 0; i < dim; i++)
        if (i != dir) acc = acc + ora_pow(s[i], 2.0, mm);
    return acc;
}
/* vectors.py:153-182; returns 0 on ValueError (math.sqrt of a negative number) */
static int disp_positive(int dim, const double *s, double nsq, int dir, int mm, double *out) {
    double a = nsq - v_others_sq(dim, s, dir, mm);
    if (a < 0.0) return 0;
    *out = s[dir] - sqrt(a);
    return 1;
}
/* vectors.py:185-214 */
static int disp_negative(int dim, const double *s, double nsq, int dir, int mm, double *out) {
    double a = nsq - v_others_sq(dim, s, dir, mm);
    if (a < 0.0) return 0;
    *out = s[dir] + sqrt(a);
    return 1;
}

/* ===== HardSpherePotential.displacement (hard_sphere_potential.py:65-99) ============== */
double ora_hs_displacement(int dim, const double *velocity, const double *separation, double radius) {
    double diameter_squared = 4.0 * radius * radius;
    double v2 = v_norm_sq(dim, velocity);
    double s2 = v_norm_sq(dim, separation);
    double vs = 0.0;
    for (int i = 0; i < dim; i++) vs = vs + velocity[i] * separation[i];
    double srt = vs * vs - v2 * (s2 - diameter_squared);
    return (srt >= 0.0 && vs >= 0.0) ? (vs - sqrt(srt)) / v2 : INFINITY;
}

/* ===== HardDipolePotential (hard_dipole_potential.py:66-110) ============================
 * NaN where the reference asserts (separation outside [minimum, maximum] by more than 1e-13,
 * zero velocity). */
double ora_hd_displacement(int dim, const double *velocity, const double *separation, double minimum,
                           double maximum) {
    const double min2 = minimum * minimum, max2 = maximum * maximum;
    const double v2 = v_norm_sq(dim, velocity);
    const double s2 = v_norm_sq(dim, separation);
    if (!(v2 > 0.0) || !(s2 - min2 > -1.0e-13) || !(max2 - s2 > -1.0e-13)) return NAN;
    double vs = 0.0;
    for (int i = 0; i < dim; i++) vs = vs + velocity[i] * separation[i];
    if (vs >= 0.0) {
        const double m = vs * vs - v2 * (s2 - min2);
        if (m >= 0.0) return (vs - sqrt(m)) / v2;
    }
    const double M = vs * vs - v2 * (s2 - max2);
    return M >= 0.0 ? (vs + sqrt(M)) / v2 : NAN;
}

/* ===== InversePowerPotential (inverse_power_potential.py:71-179) ====================== */
double ora_ip_potential(int dim, const double *s, double charge_product, double power, double prefactor, int mm) {
    return charge_product * prefactor / ora_pow(v_norm_sq(dim, s), power / 2.0, mm);
}

double ora_ip_derivative(int dim, int dir, double speed, const double *s, double c1, double c2, double power,
                         double prefactor, int mm) {
    double d = power * s[dir] / ora_pow(v_norm(dim, s, mm), power + 2, mm) * prefactor * c1 * c2;
    return d * speed;
}

double ora_ip_displacement(int dim, int dir, double speed, const double *sep, double c1, double c2, double power,
                           double prefactor, double dU, int mm) {
    double s[3];
    memcpy(s, sep, sizeof(double) * dim);
    double cp = c1 * c2;
    double pp = prefactor * cp;
    double two_over_power = 2.0 / power;
    double out;
    if (pp > 0) {
        if (s[dir] <= 0.0) return INFINITY;
        double smax[3];
        memcpy(smax, s, sizeof(double) * dim);
        smax[dir] = 0.0;
        double maxp = ora_ip_potential(dim, smax, cp, power, prefactor, mm);
        double cur = ora_ip_potential(dim, s, cp, power, prefactor, mm);
        if (dU < maxp - cur) {
            double nsq = ora_pow(cp * prefactor / (cur + dU), two_over_power, mm);
            if (!disp_positive(dim, s, nsq, dir, mm, &out)) return NAN;
            return out / speed;
        }
        return INFINITY;
    } else {
        double disp = 0.0;
        if (s[dir] > 0.0) { disp += s[dir]; s[dir] = 0.0; }
        double cur = ora_ip_potential(dim, s, cp, power, prefactor, mm);
        if (cur + dU >= 0.0) return INFINITY;
        double nsq = ora_pow(cp * prefactor / (cur + dU), two_over_power, mm);
        if (!disp_negative(dim, s, nsq, dir, mm, &out)) return NAN;
        disp += out;
        return disp / speed;
    }
}

/* Truncated-and-shifted inverse power U(r) - U(rc) for r < rc (no reference counterpart).
 * Same inversion as above with the potential energy floor U(rc) in place of zero. */
double ora_ipt_displacement(int dim, int dir, double speed, const double *sep, double c1, double c2, double power,
                            double prefactor, double cutoff, double dU, int mm) {
    double s[3];
    memcpy(s, sep, sizeof(double) * dim);
    double cp = c1 * c2;
    double pp = prefactor * cp;
    double rc2 = cutoff * cutoff;
    double perp2 = 0.0;
    for (int i = 0; i < dim; i++)
        if (i != dir) perp2 = perp2 + s[i] * s[i];
    if (perp2 >= rc2) return INFINITY;
    double two_over_power = 2.0 / power;
    double urc = cp * prefactor / ora_pow(rc2, power / 2.0, mm);
    double out;
    if (pp > 0) {
        if (s[dir] <= 0.0) return INFINITY;
        double s2 = v_norm_sq(dim, s);
        double cur = s2 < rc2 ? ora_ip_potential(dim, s, cp, power, prefactor, mm) : urc;
        double smax[3];
        memcpy(smax, s, sizeof(double) * dim);
        smax[dir] = 0.0;
        double maxp = ora_ip_potential(dim, smax, cp, power, prefactor, mm);
        if (dU < maxp - cur) {
            double nsq = ora_pow(cp * prefactor / (cur + dU), two_over_power, mm);
            double a = nsq - perp2;
            if (a < 0.0) return NAN;
            out = s[dir] - sqrt(a);
            return out / speed;
        }
        return INFINITY;
    } else {
        double disp = 0.0;
        if (s[dir] > 0.0) { disp += s[dir]; s[dir] = 0.0; }
        double s2 = v_norm_sq(dim, s);
        if (s2 >= rc2) return INFINITY;
        double cur = ora_ip_potential(dim, s, cp, power, prefactor, mm);
        if (cur + dU >= urc) return INFINITY;
        double nsq = ora_pow(cp * prefactor / (cur + dU), two_over_power, mm);
        double a = nsq - perp2;
        if (a < 0.0) return NAN;
        disp += s[dir] + sqrt(a);
        return disp / speed;
    }
}

/* ===== MexicanHatPotential (potential/abstracts.py:285-581) ============================ */
typedef struct {
    int kind;            /* 0: Lennard-Jones, 1: displaced even power */
    double prefactor, sigma, r0, r0sq;
    double k6, k12;      /* LJ inverse-power prefactors */
    int power;           /* displaced even power */
    double inverse_power;
    int dim, mm;
} mh_t;

static double mh_potential(const mh_t *h, const double *s) {
    if (h->kind == 0) {
        /* lennard_jones_potential.py:83-98: six.potential(1.0, s) + twelve.potential(1.0, s) */
        return ora_ip_potential(h->dim, s, 1.0, 6.0, h->k6, h->mm) + ora_ip_potential(h->dim, s, 1.0, 12.0, h->k12, h->mm);
    }
    /* displaced_even_power_potential.py:95-110 */
    double d = v_norm(h->dim, s, h->mm) - h->r0;
    return h->prefactor * ora_pow(d, (double)h->power, h->mm);
}
static double mh_invert_inside(const mh_t *h, double u) {
    if (h->kind == 0) {
        double so6 = (1 + ora_pow(1 + 4 * u / h->prefactor, 0.5, h->mm)) / 2;
        return h->sigma / ora_pow(so6, 1.0 / 6.0, h->mm);
    }
    return h->r0 - ora_pow(u / h->prefactor, h->inverse_power, h->mm);
}
static double mh_invert_outside(const mh_t *h, double u) {
    if (h->kind == 0) {
        if (u >= 0.0) return INFINITY;
        double so6 = (1 - ora_pow(1 + 4 * u / h->prefactor, 0.5, h->mm)) / 2;
        return h->sigma / ora_pow(so6, 1.0 / 6.0, h->mm);
    }
    return h->r0 + ora_pow(u / h->prefactor, h->inverse_power, h->mm);
}
/* Each helper returns 0 when the Python code would raise ValueError (math.sqrt < 0). */
static int mh_front_outside(const mh_t *h, int dir, double cur, double dU, double *s, double *out) {
    double nn = mh_invert_outside(h, cur + dU);
    return disp_negative(h->dim, s, nn * nn, dir, h->mm, out);
}
static int mh_front_inside(const mh_t *h, int dir, double dU, double *s, double *out) {
    double disp;
    if (!disp_negative(h->dim, s, h->r0sq, dir, h->mm, &disp)) return 0;
    s[dir] -= disp;
    double cur = mh_potential(h, s);
    double more;
    if (!mh_front_outside(h, dir, cur, dU, s, &more)) return 0;
    *out = disp + more;
    return 1;
}
static int mh_behind_inside(const mh_t *h, int dir, double cur, double dU, double *s, double *out) {
    double smax[3];
    memcpy(smax, s, sizeof(double) * h->dim);
    smax[dir] = 0.0;
    double maxp = mh_potential(h, smax);
    double diff = maxp - cur;
    if (dU < diff) {
        double nn = mh_invert_inside(h, cur + dU);
        return disp_positive(h->dim, s, nn * nn, dir, h->mm, out);
    }
    double disp = s[dir];
    s[dir] = 0.0;
    dU -= diff;
    double more;
    if (!mh_front_inside(h, dir, dU, s, &more)) return 0;
    *out = disp + more;
    return 1;
}
static int mh_behind_outside(const mh_t *h, int dir, double dU, double *s, double *out) {
    double disp, more;
    int ok = disp_positive(h->dim, s, h->r0sq, dir, h->mm, &disp);
    if (ok) {
        s[dir] -= disp;
        double cur = mh_potential(h, s);
        ok = mh_behind_inside(h, dir, cur, dU, s, &more);
        if (ok) { *out = disp + more; return 1; }
    }
    /* except ValueError: (separation as mutated so far) */
    disp = s[dir];
    s[dir] = 0.0;
    double cur = mh_potential(h, s);
    if (!mh_front_outside(h, dir, cur, dU, s, &more)) return 0;
    *out = disp + more;
    return 1;
}
static double mh_displacement(const mh_t *h, int dir, double speed, const double *sep, double dU) {
    double s[3], out;
    memcpy(s, sep, sizeof(double) * h->dim);
    double nrm = v_norm(h->dim, s, h->mm);
    int ok;
    if (nrm >= h->r0) {
        if (s[dir] <= 0.0) {
            double cur = mh_potential(h, s);
            ok = mh_front_outside(h, dir, cur, dU, s, &out);
        } else {
            ok = mh_behind_outside(h, dir, dU, s, &out);
        }
    } else {
        if (s[dir] <= 0.0) {
            ok = mh_front_inside(h, dir, dU, s, &out);
        } else {
            double cur = mh_potential(h, s);
            ok = mh_behind_inside(h, dir, cur, dU, s, &out);
        }
    }
    if (!ok) return NAN;
    return out / speed;
}

static void mh_lj(mh_t *h, int dim, double prefactor, double sigma, int mm) {
    h->kind = 0; h->dim = dim; h->mm = mm;
    h->prefactor = prefactor; h->sigma = sigma;
    h->r0 = sigma * ora_pow(2.0, 1.0 / 6.0, mm);
    h->r0sq = ora_pow(h->r0, 2.0, mm);
    h->k6 = -prefactor * ora_pow(sigma, 6.0, mm);
    h->k12 = prefactor * ora_pow(sigma, 12.0, mm);
}

double ora_lj_displacement(int dim, int dir, double speed, const double *s, double prefactor, double sigma,
                           double dU, int mm) {
    mh_t h;
    mh_lj(&h, dim, prefactor, sigma, mm);
    return mh_displacement(&h, dir, speed, s, dU);
}

double ora_lj_derivative(int dim, int dir, double speed, const double *s, double prefactor, double sigma, int mm) {
    mh_t h;
    mh_lj(&h, dim, prefactor, sigma, mm);
    /* lennard_jones_potential.py:63-81 */
    double d = ora_ip_derivative(dim, dir, 1.0, s, 1.0, 1.0, 6.0, h.k6, mm)
             + ora_ip_derivative(dim, dir, 1.0, s, 1.0, 1.0, 12.0, h.k12, mm);
    return d * speed;
}

static void mh_dep(mh_t *h, int dim, double r0, int power, double prefactor, int mm) {
    h->kind = 1; h->dim = dim; h->mm = mm;
    h->prefactor = prefactor; h->r0 = r0; h->r0sq = r0 * r0;
    h->power = power; h->inverse_power = 1.0 / power;
}

double ora_dep_displacement(int dim, int dir, double speed, const double *s, double r0, int power, double prefactor,
                            double dU, int mm) {
    mh_t h;
    mh_dep(&h, dim, r0, power, prefactor, mm);
    return mh_displacement(&h, dir, speed, s, dU);
}

double ora_dep_derivative(int dim, int dir, double speed, const double *s, double r0, int power, double prefactor,
                          int mm) {
    /* displaced_even_power_potential.py:73-93 */
    double nrm = v_norm(dim, s, mm);
    double d = -power * prefactor * ora_pow(nrm - r0, (double)(power - 1), mm) * s[dir] / nrm;
    return d * speed;
}

/* ===== InversePowerCoulombBoundingPotential (…bounding_potential.c:53-139) ============ */
double ora_ipcb_derivative(double pp, double sx, double sy, double sz) {
    return pp * sx / pow(sx * sx + sy * sy + sz * sz, 3.0 / 2.0);
}
static double ipcb_potential(double pp, double sx, double sy, double sz) {
    return pp / sqrt(sx * sx + sy * sy + sz * sz);
}
double ora_ipcb_displacement(double pp, double sx, double sy, double sz, double dU, double L) {
    double half = L / 2.0;
    double cur = ipcb_potential(pp, sx, sy, sz);
    double p0 = ipcb_potential(pp, 0.0, sy, sz);
    double ph = ipcb_potential(pp, half, sy, sz);
    double per_length = fabs(p0 - ph);
    double disp = floor(dU / per_length) * L;
    dU = fmod(dU, per_length);
    double nn;
    if (pp > 0.0) {
        if (sx <= 0.0) {
            disp += half + sx; sx = half; cur = ph;
        } else if (dU >= p0 - cur) {
            dU -= (p0 - cur); disp += sx + half; sx = half; cur = ph;
        }
        nn = pp / (cur + dU);
        disp += (sx - sqrt(nn * nn - (sy * sy + sz * sz)));
    } else {
        if (sx > 0.0) {
            disp += sx; sx = 0.0; cur = p0;
        } else if (dU >= ph - cur) {
            dU -= (ph - cur); disp += sx + L; sx = 0.0; cur = p0;
        }
        nn = pp / (cur + dU);
        disp += (sx + sqrt(nn * nn - (sy * sy + sz * sz)));
    }
    return disp;
}

/* ===== MergedImageCoulombPotential (merged_image_coulomb_potential.c:77-274) ========== */
struct ora_ewald {
    int fc, fc2, pc, pc2;
    double aol, aol2, two_aol_rpi, L, tpol;
    double *f;  /* (fc+1)^3, index [i][j][k] */
};
ora_ewald *ora_ewald_create(int fc, int pc, double alpha, double L) {
    ora_ewald *e = (ora_ewald *)calloc(1, sizeof(ora_ewald));
    if (!e) return NULL;
    int n = fc + 1;
    e->f = (double *)calloc((size_t)n * n * n, sizeof(double));
    if (!e->f) { free(e); return NULL; }
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 1; i < n; i++) {
                double coeff = (j == 0 && k == 0) ? 1.0 : ((k == 0 || j == 0) ? 2.0 : 4.0);
                double nsq = i * i + j * j + k * k;
                e->f[(i * n + j) * n + k] = 4.0 * i * coeff / (nsq * L * L) * exp(-M_PI * M_PI * nsq / (alpha * alpha));
            }
    e->fc = fc; e->fc2 = fc * fc; e->pc = pc; e->pc2 = pc * pc;
    e->aol = alpha / L; e->aol2 = alpha * alpha / (L * L);
    e->two_aol_rpi = 2.0 * alpha / (L * sqrt(M_PI));
    e->L = L; e->tpol = 2.0 * M_PI / L;
    return e;
}
void ora_ewald_destroy(ora_ewald *e) { if (e) { free(e->f); free(e); } }

double ora_ewald_derivative(const ora_ewald *e, double sx, double sy, double sz) {
    double der = 0.0;
    int n = e->fc + 1;
    for (int k = -e->pc; k < e->pc + 1; k++) {
        double vz2 = (sz + k * e->L) * (sz + k * e->L);
        int cy = (int)sqrt(e->pc2 - k * k);
        for (int j = -cy; j < cy + 1; j++) {
            double vy2 = (sy + j * e->L) * (sy + j * e->L);
            int cx = (int)sqrt(e->pc2 - j * j - k * k);
            for (int i = -cx; i < cx + 1; i++) {
                double vx = sx + i * e->L;
                double v2 = vx * vx + vy2 + vz2;
                double vn = sqrt(v2);
                der += vx * (e->two_aol_rpi * exp(-e->aol2 * v2) + erfc(e->aol * vn) / vn) / v2;
            }
        }
    }
    double dcx = cos(e->tpol * sx), dsx = sin(e->tpol * sx);
    double dcy = cos(e->tpol * sy), dsy = sin(e->tpol * sy);
    double dcz = cos(e->tpol * sz), dsz = sin(e->tpol * sz);
    double cx_ = dcx, sx_ = dsx, cy_ = 1.0, sy_ = 0.0, cz_ = 1.0, sz_ = 0.0, t;
    for (int i = 1; i < e->fc + 1; i++) {
        int cy = (int)sqrt(e->fc2 - i * i);
        for (int j = 0; j < cy + 1; j++) {
            int cx = (int)sqrt(e->fc2 - i * i - j * j);
            for (int k = 0; k < cx + 1; k++) {
                der += e->f[(i * n + j) * n + k] * sx_ * cy_ * cz_;
                if (k != cx) {
                    t = cz_; cz_ = t * dcz - sz_ * dsz; sz_ = sz_ * dcz + t * dsz;
                } else if (j != cy) {
                    t = cy_; cy_ = t * dcy - sy_ * dsy; sy_ = sy_ * dcy + t * dsy;
                    cz_ = 1.0; sz_ = 0.0;
                } else if (i != e->fc) {
                    t = cx_; cx_ = t * dcx - sx_ * dsx; sx_ = sx_ * dcx + t * dsx;
                    cy_ = 1.0; sy_ = 0.0; cz_ = 1.0; sz_ = 0.0;
                }
            }
        }
    }
    return der;
}
