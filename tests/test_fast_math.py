"""The fast paths of the correctly rounded log and pow (jellyfysh_amd/csrc/ecmc_math.h), built
for the host from the same header the kernels include, against the oracle's quad-precision
rounding (oracle/ecmc_oracle_math.c ora_log / ora_pow in correctly rounded mode).

Pins: every result equals the correctly rounded one; the fast log's own error (hi + lo
against quad precision) stays below the 2^-64 its rounding test assumes, the k-th root path's
and the half-integer power path's below 2^-92, the Newton path for 2/p below 2^-88; the
rounding test sends only a small share of calls to the slow path; the double-double x^3 and
x^6 of the inversion round correctly too."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#define ECMC_HOST_BUILD 1
#include <stdio.h>
#include <stdlib.h>
#include <quadmath.h>
#include "ecmc_math.h"
extern "C" double ora_pow(double x, double y, int math_mode);
extern "C" double ora_log(double x, int math_mode);
using namespace ecmc;
static unsigned long long st = 88172645463325252ull;
static double unif() {                       /* xorshift64, 53-bit uniform in [0, 1) */
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return (double)(st >> 11) * 0x1p-53;
}
int main(int argc, char **argv) {
    long n = atol(argv[1]), bad_log = 0, bad_pow = 0, slow_log = 0, slow_pow = 0, slow_half = 0, slow_rat = 0;
    __float128 worst_log = 0, worst_pow = 0, worst_half = 0, worst_rat = 0;
    for (long i = 0; i < n; i++) {
        double x;
        switch (i % 5) {
        case 0: x = 1.0 - unif(); break;
        case 1: x = exp(unif() * 80.0 - 40.0); break;
        case 2: x = 1.0 + (unif() - 0.5) * 0x1p-6; break;
        case 3: x = 1.0 + (unif() - 0.5) * 0x1p-30; break;
        default: x = ldexp(1.0 + unif(), (int)(unif() * 2000) - 1000); break;
        }
        if (!(x > 0.0) || x == 1.0) continue;
        double lo, hi = log_fast(x, lo);
        __float128 t = logq((__float128)x), e = fabsq(((__float128)hi + lo - t) / t);
        if (e > worst_log) worst_log = e;
        slow_log += !rounding_settled(hi, lo, 0x1p-64);
        bad_log += cr_log(x) != ora_log(x, 1);
        double y = (i & 1) ? 2.0 / 12.0 : 2.0 / 6.0, b = exp(unif() * 30.0 - 8.0);
        double plo, phi = root_pow_fast(b, y == 1.0 / 6.0, plo);
        __float128 tp = powq((__float128)b, (__float128)y), ep = fabsq(((__float128)phi + plo - tp) / tp);
        if (ep > worst_pow) worst_pow = ep;
        slow_pow += !rounding_settled(phi, plo, 0x1p-92);
        bad_pow += cr_pow(b, y) != ora_pow(b, y, 1);
        bad_pow += root36(b, y == 1.0 / 6.0) != ora_pow(b, y, 1);
        /* the integer powers of the inversion: (r^2)^6 and (r^2)^3 */
        double r2 = exp(unif() * 6.0 - 3.0);
        bad_pow += pow36(r2, true) != ora_pow(r2, 6.0, 1);
        bad_pow += pow36(r2, false) != ora_pow(r2, 3.0, 1);
        /* half-integer powers (r^2)^(p/2) of odd p, and 2/p of other integer powers */
        static const double hy[4] = {4.5, 3.5, 1.5, 15.5};
        double hx = exp(unif() * 40.0 - 20.0), hyv = hy[i % 4];
        double hlo, hhi = half_pow_fast(hx, (int)(hyv - 0.5), hlo);
        __float128 th = powq((__float128)hx, (__float128)hyv), eh = fabsq(((__float128)hhi + hlo - th) / th);
        if (eh > worst_half) worst_half = eh;
        slow_half += !rounding_settled(hhi, hlo, 0x1p-92);
        bad_pow += cr_pow(hx, hyv) != ora_pow(hx, hyv, 1);
        static const int pp[6] = {9, 7, 5, 10, 16, 31};
        const int pv = pp[i % 6];
        const double ry = 2.0 / pv, rx = exp(unif() * 60.0 - 20.0);
        double rlo, rhi = rational_pow_fast(rx, ry, (pv & 1) ? 2 : 1, (pv & 1) ? pv : pv / 2, rlo);
        __float128 tr = powq((__float128)rx, (__float128)ry), er = fabsq(((__float128)rhi + rlo - tr) / tr);
        if (er > worst_rat) worst_rat = er;
        slow_rat += !rounding_settled(rhi, rlo, 0x1p-88);
        bad_pow += cr_pow(rx, ry) != ora_pow(rx, ry, 1);
    }
    printf("%ld %ld %ld %ld %.3f %.3f %ld %ld %.3f %.3f\n", bad_log, bad_pow, slow_log, slow_pow,
           (double)log2q(worst_log), (double)log2q(worst_pow), slow_half, slow_rat, (double)log2q(worst_half),
           (double)log2q(worst_rat));
    return 0;
}
"""


def test_fast_log_and_root_pow_round_correctly(tmp_path):
    gxx = shutil.which("g++")
    oracle_lib = os.path.join(ROOT, "oracle", "_build", "libecmc_oracle.so")
    if gxx is None or not os.path.exists(oracle_lib):
        pytest.skip("needs g++ and the built oracle")
    src = tmp_path / "fast_math.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "fast_math"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "jellyfysh_amd", "csrc"),
                    str(src), "-o", str(exe), oracle_lib, "-Wl,-rpath," + os.path.dirname(oracle_lib),
                    "-lquadmath"], check=True, capture_output=True)
    n = 400_000
    out = subprocess.run([str(exe), str(n)], capture_output=True, text=True, check=True).stdout.split()
    bad_log, bad_pow, slow_log, slow_pow = (int(v) for v in out[:4])
    worst_log, worst_pow = float(out[4]), float(out[5])
    slow_half, slow_rat = int(out[6]), int(out[7])
    worst_half, worst_rat = float(out[8]), float(out[9])
    assert bad_log == 0 and bad_pow == 0
    assert worst_log < -64.0 and worst_pow < -92.0
    assert worst_half < -92.0 and worst_rat < -88.0        # the accuracies their rounding tests assume
    assert slow_log < 0.005 * n and slow_pow == 0 and slow_half == 0 and slow_rat == 0


MODPROG = r"""
#define ECMC_HOST_BUILD 1
#include <stdio.h>
#include "ecmc_math.h"
using namespace ecmc;
int main() {
    double x, y;
    while (scanf("%la %la", &x, &y) == 2) {
        double q, r;
        py_divmod(x, y, q, r);
        printf("%a %a %a %a\n", py_mod(x, y), q, r, fmod_near(fabs(x), y));
    }
    return 0;
}
"""


def test_python_float_mod_and_divmod_of_the_kernels(tmp_path):
    """py_mod and py_divmod of ecmc_math.h (the kernels' separation wrap and Time's divmod by one,
    with their one-period and non-negative fast paths) equal CPython's % and divmod, and
    fmod_near (the image windings of the Coulomb bound) equals math.fmod, signed zeros included,
    around every branch point: 0, +-y, +-2y, and far outside."""
    import math
    import random
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("needs g++")
    src = tmp_path / "mod.cpp"
    src.write_text(MODPROG)
    exe = tmp_path / "mod"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "jellyfysh_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True, capture_output=True)
    rnd = random.Random(5)
    pairs = []
    for y in (1.0, 0.5, 17.2355, 8.0, 2.0 ** -3, 3.0, 1e-300):
        edges = [0.0, -0.0, y, -y, 2 * y, -2 * y, 0.5 * y, -0.5 * y]
        for e in edges:
            for k in range(-3, 4):
                v = e
                for _ in range(abs(k)):
                    v = math.nextafter(v, math.inf if k > 0 else -math.inf)
                pairs.append((v, y))
        for _ in range(3000):
            pairs.append((rnd.uniform(-2.5 * y, 2.5 * y), y))
        pairs += [(rnd.uniform(-1e6, 1e6) * y, y) for _ in range(200)]
        pairs += [(1e300, y), (-1e300, y), (5e-324, y), (-5e-324, y)]
    inp = "".join("%s %s\n" % (x.hex(), y.hex()) for x, y in pairs)
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for (x, y), line in zip(pairs, out):
        m, q, r, f = (float.fromhex(v) for v in line.split())
        wq, wr = divmod(x, y)
        for got, want in ((m, x % y), (q, wq), (r, wr), (f, math.fmod(abs(x), y))):
            assert got == want and math.copysign(1, got) == math.copysign(1, want), (x, y, got, want)
