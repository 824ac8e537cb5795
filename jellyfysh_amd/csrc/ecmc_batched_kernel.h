// ecmc_batched_kernel.h -- device code of the batched replicas (see ecmc_batched.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/jellyfysh_amd.h"
#include "ecmc_math.h"
#include "ecmc_potentials.h"
#include "ecmc_fp32_bounds.h"

namespace ecmc {

constexpr int B_MAX_WPR = 8;
constexpr int B_CONTENDERS = 48;     // contenders per leg (about 7 typical; beyond, all candidates go exact)
constexpr int B_QUEUE = 128;         // per-wavefront queue of candidates past the geometric test
constexpr int B_MAX_CPS = 8;         // cells per side handled in LDS (512 cells)
// the chunk start counts units per cell in the (then idle) contender and queue area
static_assert(36 * B_CONTENDERS + 4 * B_QUEUE >= 4 * B_MAX_CPS * B_MAX_CPS * B_MAX_CPS, "cell counts do not fit");

struct ReplicaDev {
    int32_t n, cps;
    int64_t off;           // offset of the replica's SoA rows in pos64 (3 * n doubles)
    double L, cutoff, side;
    int32_t bound_off, pad;
    // state
    int32_t active, direction;
    double chain_left;
    uint64_t leg;
    double est[8];
    int32_t err, pad2;
    int64_t stats[4];      // candidates, geometric survivors, contenders, exact evaluations
};

struct FactorDev {
    int32_t n_factors, pad;
    double power[2], prefactor[2];
};

struct BatchArgs {
    ReplicaDev *rep;
    double *pos64;
    const double *bounds;        // cell lower bounds, per replica cps+1 entries
    const int32_t *order;        // replica indices of this class, largest first
    int32_t *state;              // per class member: (chunks done << 1) | held
    int32_t count;               // replicas in this class
    int32_t nmax, cellmax;       // LDS sizing of this class
    int32_t legs;                // legs per replica of this step
    int32_t chunk_legs, nchunks; // a work item is one chunk of legs of one replica
    int32_t mode;                // 0 filtered, 1 all candidates exact, 2 filtered + audit
    int32_t log_legs;            // legs logged per replica (0 = none)
    uint32_t seed_lo, pad;
    double beta, chain_length;
    FactorDev fac;
    int32_t *queue;              // [0] claim cursor
    ecmc_batch_event *log;       // n_replicas x log_legs
    unsigned long long *audit;   // bound violations
};

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) --------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ double unit53(uint32_t a, uint32_t b) {
    return ((a >> 5) * 67108864.0 + (b >> 6)) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double bmin_image(double s, double L) {
    if (s >= 0.5 * L) return s - L;
    if (s < -0.5 * L) return s + L;
    return s;
}

__device__ __forceinline__ int bcell1(double x, double side, int cps) {
    int c = (int)(x / side);
    return c < cps ? c : cps - 1;
}

// ---- FP32 lower bound of the truncated inverse-power event displacement ------------------
// (the bracketing helpers up, dn, rcp1, sqrt_up, sqrt_dn and dU_lower are in ecmc_fp32_bounds.h)

// V(r2) = K / r2^(p/2) with p in {6, 12} (integer powers by multiplication), bounded up
// (hi) or down for an r2 that is itself a lower/upper bound.  Three products (2^-24 each) and
// K * rcp(rp) (3 x 2^-24) stay inside an 8 x 2^-24 allowance, to which ETA is added; the two
// factors are one constant, (1 +- 2^-21)(1 +- 2^-17) without its 2^-38 term (the side that
// widens the bound), applied with one more rounding.
__device__ __forceinline__ float vpow_bound(float K, float r2, int half_p, bool hi) {
    float r3 = r2 * r2 * r2;
    float rp = half_p == 6 ? r3 * r3 : r3;
    float v = K * rcp1(rp);
    return v * (hi ? 1.f + 0x1p-17f + 0x1p-21f : 1.f - 0x1p-17f - 0x1p-21f);
}

// (K / V)^(2/p) bounded up (hi) or down.
__device__ __forceinline__ float rstar2_bound(float K, float V, float two_over_p, bool hi) {
    float q = K * rcp1(V);
    float r = exp2f(two_over_p * log2f(q));
    return hi ? r * (1.f + ETA_POW) : r * (1.f - ETA_POW);
}

struct Geo {
    float sd_lo, sd_hi, perp2_lo, perp2_hi, r2_lo, r2_hi;
};

// Lower bound of the exact displacement of one factor, or +INF when the factor surely has
// no event, given the exponential draw's lower bound dU_lo.
__device__ float factor_lower_bound(const Geo &g, float k, int p, float rc2_hi, float vrc_lo, float dU_lo) {
    const float INF = __int_as_float(0x7f800000);
    const int half_p = p / 2;
    const float two_over_p = half_p == 6 ? 1.f / 6.f : 1.f / 3.f;
    if (g.perp2_lo >= rc2_hi) return INF;
    if (k > 0.f) {
        if (g.sd_hi <= 0.f) return INF;
        float ueff_lo = fmaxf(vpow_bound(k, g.r2_hi, half_p, false), vrc_lo);
        float umax_hi = vpow_bound(k, fmaxf(g.perp2_lo, 1e-30f), half_p, true);
        float room_hi = up(umax_hi - ueff_lo);
        if (dU_lo >= room_hi) return INF;
        float ut_lo = dn(ueff_lo + dU_lo);
        float rs2_hi = rstar2_bound(k, ut_lo, two_over_p, true);
        float x_hi = sqrt_up(fmaxf(up(rs2_hi - g.perp2_lo), 0.f));
        float d = dn(g.sd_lo) - x_hi;
        d = d - fabsf(d) * 0x1p-22f;
        return fmaxf(d, 0.f);
    }
    const float K = -k;
    float x0sq_lo = g.sd_hi < 0.f ? sq_lo(-g.sd_hi) : 0.f;
    float r02_lo = dn(g.perp2_lo + x0sq_lo);
    if (r02_lo >= rc2_hi) return INF;
    float vcur_hi = vpow_bound(K, fmaxf(r02_lo, 1e-30f), half_p, true);
    float room_hi = up(vcur_hi - vrc_lo);
    if (dU_lo >= room_hi) return INF;
    float vt_hi = up(vcur_hi - dU_lo);
    if (!(vt_hi > 0.f)) return 0.f;
    float rs2_lo = rstar2_bound(K, vt_hi, two_over_p, false);
    float y_lo = sqrt_dn(fmaxf(dn(rs2_lo - g.perp2_hi), 0.f));
    float d = g.sd_lo + y_lo;
    d = d - fabsf(d) * 0x1p-22f - 0x1p-40f;
    return fmaxf(d, 0.f);
}

// Smallest displacement any event of the factor can have, from geometry alone: a repulsive
// factor first has to reach the cutoff sphere, an attractive one first passes the closest
// approach.
__device__ __forceinline__ float factor_geometric_floor(const Geo &g, float k, float rc2_hi) {
    if (k > 0.f) {
        if (g.r2_lo < rc2_hi) return 0.f;
        float entry_hi = sqrt_up(fmaxf(up(rc2_hi - g.perp2_lo), 0.f));
        float d = dn(g.sd_lo) - entry_hi;
        d = d - fabsf(d) * 0x1p-22f;
        return fmaxf(d, 0.f);
    }
    return fmaxf(dn(g.sd_lo), 0.f);
}

struct BKey {
    double d;
    int32_t rank, j, f;
};
__device__ __forceinline__ bool bkey_lt(const BKey &a, const BKey &b) {
    if (a.d != b.d) return a.d < b.d;
    if (a.rank != b.rank) return a.rank < b.rank;
    if (a.j != b.j) return a.j < b.j;
    return a.f < b.f;
}
__device__ __forceinline__ bool bkey_eq(const BKey &a, const BKey &b) {
    return a.d == b.d && a.rank == b.rank && a.j == b.j && a.f == b.f;
}

// One step of the (displacement, key) minimum across lanes with a DPP lane permutation of
// the 16-lane row (CTRL: 0xB1 swaps neighbours, 0x4E pairs of neighbours, 0x141 mirrors each
// half row, 0x140 the row), in registers.
template <int CTRL>
__device__ __forceinline__ void dpp_min_step(double &d, int &k) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(d), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(d), CTRL, 0xF, 0xF, false);
    const int ok = __builtin_amdgcn_update_dpp(0, k, CTRL, 0xF, 0xF, false);
    const double od = __hiloint2double(hi, lo);
    if (od < d || (od == d && ok < k)) { d = od; k = ok; }
}

// The leg's event over the workgroup: the minimum over lanes of (displacement, packed (kind,
// target, factor)), then LDS across wavefronts (one barrier).  With one wavefront per replica
// the first 16 lanes reduce by DPP and the other rows join only when more than 16 lanes hold
// keys (nlive: lanes beyond hold +inf), the result taken from lane 0, and the returned index
// is the lane holding the winning key (keys are unique; 0 unless a pair wins), from which the
// caller reads the candidate's position; with several, a shuffle butterfly, the winning lane
// of each wavefront leaves that candidate's FP64 position in newpos[3 * wave], and the
// returned index is the winning wavefront.
template <int WPR>
__device__ int block_min(BKey &b, const BKey &mine, const double *pos, BKey *sk, double *newpos, int nlive) {
    // (rank, j, f) packed into one word orders like the triple (rank < 4, j < 2^27, f < 2)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double d = b.d;
    int k = (b.rank << 28) | (b.j << 1) | b.f;
    if (WPR == 1) {
        dpp_min_step<0xB1>(d, k);
        dpp_min_step<0x4E>(d, k);
        dpp_min_step<0x141>(d, k);
        dpp_min_step<0x140>(d, k);
        if (nlive > 16) {
#pragma unroll
            for (int off = 16; off <= 32; off <<= 1) {
                const double od = __shfl_xor(d, off, 64);
                const int ok = __shfl_xor(k, off, 64);
                if (od < d || (od == d && ok < k)) { d = od; k = ok; }
            }
        }
        d = __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(d)),
                             __builtin_amdgcn_readfirstlane(__double2loint(d)));
        k = __builtin_amdgcn_readfirstlane(k);
        b.d = d;
        b.rank = k >> 28;
        b.j = (k >> 1) & 0x7ffffff;
        b.f = k & 1;
        const unsigned long long wm = __ballot(b.rank == 0 && bkey_eq(mine, b));
        return wm ? __ffsll((long long)wm) - 1 : 0;
    } else {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double od = __shfl_xor(d, off, 64);
            const int ok = __shfl_xor(k, off, 64);
            if (od < d || (od == d && ok < k)) { d = od; k = ok; }
        }
    }
    b.d = d;
    b.rank = k >> 28;
    b.j = (k >> 1) & 0x7ffffff;
    b.f = k & 1;
    if (b.rank == 0 && bkey_eq(mine, b)) {
        newpos[3 * wave] = pos[0];
        newpos[3 * wave + 1] = pos[1];
        newpos[3 * wave + 2] = pos[2];
    }
    int wb = 0;
    if (WPR > 1) {
        if (lane == 0) sk[wave] = b;
        __syncthreads();
        b = sk[0];
#pragma unroll
        for (int w = 1; w < WPR; w++)
            if (bkey_lt(sk[w], b)) {
                b = sk[w];
                wb = w;
            }
    }
    return wb;
}

__device__ __forceinline__ void bsync() { __syncthreads(); }

// Factor-indexed values held in registers (a runtime index into a private array would put
// the array in scratch memory).
template <class T>
__device__ __forceinline__ T pick(int f, T v0, T v1) { return f ? v1 : v0; }

// Exact FP64 event displacement of one truncated-and-shifted inverse-power factor
// (k / r^p - k / rc^p inside rc), p = 12 (six: r^p = (r^2)^6) or 6, for separation s (target
// minus active, minimum image) along +d with unit speed: the values of ipt_displacement (and
// so of the oracle), each pow correctly rounded.  The repulsive (k > 0) and attractive forms
// run as one instruction stream with selects, so lanes of a wavefront evaluating different
// factors do not diverge; sum(v ** 2 for v in others) is also the maximum potential's
// squared distance (the d component set to zero adds an exact 0).  urc = k / rc^p.
#ifndef ECMC_BATCHED_WAVES_PER_SIMD
#define ECMC_BATCHED_WAVES_PER_SIMD 2   // waves per SIMD the register allocation must allow (VGPRs <= 256)
#endif
__device__ __forceinline__ double exact_ipt(double s0, double s1, double s2, int d, bool six, double prefactor,
                                            double rc2, double urc, double dU) {
    const double perp2 = d == 0 ? s1 * s1 + s2 * s2 : (d == 1 ? s0 * s0 + s2 * s2 : s0 * s0 + s1 * s1);
    const bool rep = prefactor > 0.0;
    const double sd = d == 0 ? s0 : (d == 1 ? s1 : s2);
    // attractive: a unit ahead is first passed (its closest approach is at s[d] = 0)
    const bool shift = !rep && sd > 0.0;
    const double disp = shift ? sd : 0.0;
    const double sdd = shift ? 0.0 : sd;
    const double c0 = d == 0 ? sdd : s0, c1 = d == 1 ? sdd : s1, c2 = d == 2 ? sdd : s2;
    const double sn2 = (0.0 + c0 * c0 + c1 * c1) + c2 * c2;
    bool live = perp2 < rc2 && (rep ? sd > 0.0 : sn2 < rc2);
    const double cur = (rep && !(sn2 < rc2)) ? urc : prefactor / pow36(live ? sn2 : 1.0, six);
    const double maxp = prefactor / pow36(rep && live ? perp2 : 1.0, six);
    live = live && (rep ? dU < maxp - cur : !(cur + dU >= urc));
    const double nsq = root36(live ? prefactor / (cur + dU) : 1.0, six);
    const double a = nsq - perp2;
    const double sq = sqrt(fmax(a, 0.0));
    const double res = a < 0.0 ? NAN : (rep ? sdd - sq : disp + (sdd + sq));
    return live ? res : INFINITY;
}

// The exact FP64 event displacement of one contender: the exponential draw dU = -log(1 - u) /
// beta from its two words (correctly rounded log), then exact_ipt.  A call, not inlined: the
// leg loop keeps its state in registers, and inlining this chain pushes the kernel past 256
// VGPRs into spills (measured slower than the call).
#ifndef ECMC_BATCHED_EXACT_ATTR
#define ECMC_BATCHED_EXACT_ATTR __noinline__
#endif
__device__ ECMC_BATCHED_EXACT_ATTR double exact_event(uint32_t wa, uint32_t wb, double beta, double s0, double s1,
                                                      double s2, int d, bool six, double prefactor, double rc2,
                                                      double urc) {
    const double dU = -cr_log(1.0 - unit53(wa, wb)) / beta;
    return exact_ipt(s0, s1, s2, d, six, prefactor, rc2, urc, dU);
}

// Positions in LDS: 16-bit fixed point per coordinate, q = round(x 2^16 / L) mod 2^16 (x in
// [0, L)), one uint16 array per axis.  The separation q_j - q_a wrapped into [-2^15, 2^15) is
// the minimum image exactly in integers; times L 2^-16 it is within L 2^-16 (two roundings of
// L 2^-17) plus L 2^-25 (the float product) of the true minimum-image separation, unless that
// is within the same distance of +-L/2.  The slack E = L 2^-15 covers both.
__device__ __forceinline__ uint16_t quant1(double x, double inv_step) { return (uint16_t)(uint32_t)(x * inv_step + 0.5); }
__device__ __forceinline__ float qsep(uint32_t qj, uint32_t qa, float step) {
    return (float)((int)((qj - qa + 0x8000u) & 0xFFFFu) - 0x8000) * step;
}

// FP32 geometry of a candidate from its minimum-image separation: bounds with the position
// slack E, and whether the minimum image is ambiguous (then the candidate is forced into the
// contender list).
__device__ __forceinline__ bool cand_geometry(float dx, float dy, float dz, int d, float Lf, float E, bool check_amb,
                                              Geo &g) {
    float sv[3] = {dx, dy, dz};
    bool ambiguous = false;
    if (check_amb) {
#pragma unroll
        for (int q = 0; q < 3; q++)
            if (fabsf(fabsf(sv[q]) - 0.5f * Lf) <= 2.f * E) ambiguous = true;
    }
    const float sd = d == 0 ? sv[0] : (d == 1 ? sv[1] : sv[2]);
    g.sd_lo = sd - E;
    g.sd_hi = sd + E;
    float p2lo = 0.f, p2hi = 0.f;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        if (q == d) continue;
        const float m = fabsf(sv[q]);
        const float lo = fmaxf(m - E, 0.f), hi = m + E;
        p2lo += lo * lo;
        p2hi += hi * hi;
    }
    g.perp2_lo = dn(p2lo);
    g.perp2_hi = up(p2hi);
    const float mdl = fmaxf(fabsf(sd) - E, 0.f), mdh = fabsf(sd) + E;
    g.r2_lo = dn(p2lo + mdl * mdl);
    g.r2_hi = up(p2hi + mdh * mdh);
    return ambiguous;
}

// Pre-filter of pass 1 on the fixed-point separation (dx, dy, dz): false only for a candidate
// that no factor can reach within the horizon h.  Every candidate with an event at most h
// away lies within rc of the segment [0, h] along the axis of motion: a repulsive factor has
// no event before its unit enters the cutoff sphere (ahead), an attractive one none for a
// unit behind beyond the cutoff (its distance only grows) and none ahead before the closest
// approach (at s_d <= h).  The fixed-point separation is within E of the true one on each
// axis, so within sqrt(3) E < 2E of it; hf >= h; RB2 = ((rc + 2E)(1 + 2^-20))^2 (1 + 2^-20)
// covers these and the FP32 roundings.  A separation within 2E of +-L/2 on any axis (minimum
// image not certain) always passes; check_amb false (uniform) skips that test where no such
// separation can lie in the capsule.
__device__ __forceinline__ bool cand_near(float dx, float dy, float dz, int d, float halfL, float E2, float RB2,
                                          float hf, bool check_amb) {
    bool amb = false;
    if (check_amb)
        amb = fabsf(fabsf(dx) - halfL) <= E2 || fabsf(fabsf(dy) - halfL) <= E2 || fabsf(fabsf(dz) - halfL) <= E2;
    const float sd = d == 0 ? dx : (d == 1 ? dy : dz);
    const float u = d == 0 ? dy : dx, v = d == 2 ? dy : dz;
    const float e = sd - fminf(fmaxf(sd, 0.f), hf);
    return amb || u * u + v * v + e * e < RB2;
}

// Occupant lists sorted by cell: move unit x, at slot p of cell X's segment, to cell Y != X
// (one wavefront, uniform arguments).  Going up (Y > X), x leaves through X's last slot (X's
// last unit fills p), every non-empty cell between passes its last unit to the slot before its
// first, and x becomes Y's first; going down is the mirror image.  Starts of the cells
// between, and of the far end, shift by one.  Each pass reads all its slots before it writes.
__device__ void csr_move(int16_t *cu, int16_t *cs, int p, int X, int Y, int x, int lane) {
    if (Y > X) {
        const int eX = cs[X + 1] - 1;
        const int vX = cu[eX];
        if (lane == 0 && p != eX) cu[p] = (int16_t)vX;
        for (int c0 = X + 1; c0 < Y; c0 += 64) {
            const int c = c0 + lane;
            int v = 0, dst = -1;
            if (c < Y) {
                const int b = cs[c], e = cs[c + 1];
                if (e > b) {
                    v = cu[e - 1];
                    dst = b - 1;
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (dst >= 0) cu[dst] = (int16_t)v;
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) cu[cs[Y] - 1] = (int16_t)x;
        __builtin_amdgcn_wave_barrier();
        for (int c = X + 1 + lane; c <= Y; c += 64) cs[c] = (int16_t)(cs[c] - 1);
    } else {
        const int fX = cs[X];
        const int vX = cu[fX];
        if (lane == 0 && p != fX) cu[p] = (int16_t)vX;
        // downwards, chunks from the top so that a chunk reads its slots before the chunk below
        // writes into them
        for (int c0 = X - 1; c0 > Y; c0 -= 64) {
            const int c = c0 - lane;
            int v = 0, dst = -1;
            if (c > Y) {
                const int b = cs[c], e = cs[c + 1];
                if (e > b) {
                    v = cu[b];
                    dst = e;
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (dst >= 0) cu[dst] = (int16_t)v;
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) cu[cs[Y + 1]] = (int16_t)x;
        __builtin_amdgcn_wave_barrier();
        for (int c = Y + 1 + lane; c <= X; c += 64) cs[c] = (int16_t)(cs[c] + 1);
    }
}

// candidate idx -> unit (binary search over the 27 prefix sums, or all units but `a`)
__device__ __forceinline__ int candidate_unit(int idx, int cps, int a, const int *pref, const int *start,
                                              const int16_t *cu) {
    if (!cps) return idx < a ? idx : idx + 1;
    int lo = 0, hi = 27;  // pref[lo] <= idx < pref[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (pref[mid] <= idx) lo = mid;
        else hi = mid;
    }
    return cu[start[lo] + (idx - pref[lo])];
}

// Three candidates at once (indices i0, i1, i2 < M with cells): a fixed five-step search over
// the 28 prefix sums (largest lo with pref[lo] <= idx, pref[27] = M), branch-free, so that the
// three chains of LDS reads interleave.
__device__ __forceinline__ void candidate_units3(int i0, int i1, int i2, const int *pref, const int *start,
                                                 const int16_t *cu, int &j0, int &j1, int &j2) {
    int lo0 = 0, lo1 = 0, lo2 = 0;
#pragma unroll
    for (int st = 16; st > 0; st >>= 1) {
        const int t0 = min(lo0 + st, 27), t1 = min(lo1 + st, 27), t2 = min(lo2 + st, 27);
        const int p0 = pref[t0], p1 = pref[t1], p2 = pref[t2];
        lo0 = p0 <= i0 ? t0 : lo0;
        lo1 = p1 <= i1 ? t1 : lo1;
        lo2 = p2 <= i2 ? t2 : lo2;
    }
    j0 = cu[start[lo0] + (i0 - pref[lo0])];
    j1 = cu[start[lo1] + (i1 - pref[lo1])];
    j2 = cu[start[lo2] + (i2 - pref[lo2])];
}

// LDS bytes of the dynamic part for a class (fixed-point positions, cell lists, contenders and
// the per-wavefront candidate queues); the host sizes launches with the same function.
__host__ __device__ inline size_t batched_lds_bytes(int nmax, int cellmax, int wpr) {
    size_t b = 8ull * nmax + 2ull * (cellmax + 1);
    b = (b + 7) & ~7ull;
    b += (24ull + 8ull + 4ull) * B_CONTENDERS;     // positions, random words, identifiers
    b += 4ull * B_QUEUE * wpr;
    return (b + 15) & ~15ull;
}

// One leg of a replica, for the workgroup that owns it:
//   1. horizon h = min(distance to the next cell wall, rest of the chain) (FP64);
//   2. pass 1a over every candidate: the fixed-point separation against the capsule of
//      radius rc about the segment of motion (cand_near); candidates outside are done, the
//      others are appended to the wavefront's queue (ballot compaction);
//   3. pass 1b over each full queue: FP32 geometry with the position slack, Philox draws,
//      FP32 lower bounds of the event displacements (FP32 dU bound from the draw's leading
//      bits); a factor whose bound is <= h is a contender;
//   4. pass 2, one contender per thread: its FP64 position from global memory, exact FP64
//      inversion with correctly rounded pow/log, workgroup min-reduce over (displacement,
//      kind, target, factor);
//   5. thread 0 commits.
template <int WPR>
__global__ __launch_bounds__(64 * WPR, ECMC_BATCHED_WAVES_PER_SIMD) void batched_kernel(BatchArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NTH = 64 * WPR;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint16_t *qx = (uint16_t *)smem, *qy = qx + A.nmax, *qz = qy + A.nmax;
    int16_t *cu = (int16_t *)(qz + A.nmax);    // units sorted by cell, the active unit left out
    int16_t *cs = cu + A.nmax;                 // cell c holds cu[cs[c] .. cs[c + 1])
    size_t off = (8ull * A.nmax + 2ull * (A.cellmax + 1) + 7) & ~7ull;
    int32_t *ccount = (int32_t *)(smem + off);                        // 24 B x B_CONTENDERS: chunk-start cell counts
    uint32_t *cont_w = (uint32_t *)(smem + off + 24 * B_CONTENDERS);  // B_CONTENDERS x 2
    int32_t *cont_id = (int32_t *)(cont_w + 2 * B_CONTENDERS);       // (j << 1) | f
    int32_t *queue = cont_id + B_CONTENDERS + wave * B_QUEUE;
    __shared__ BKey sk[B_MAX_WPR];
    __shared__ int s_item, s_chunk, s_nc, s_overflow, s_nearby[27], s_pref[28], s_M, s_Mfull;
    __shared__ int s_preff[28];   // audit mode: prefix sums over every unit of the 27 cells (none culled)
    __shared__ double s_ap[3], s_newpos[3 * B_MAX_WPR], s_h, s_cb, s_bounds[B_MAX_CPS + 1];
    __shared__ ReplicaDev R;
    __shared__ unsigned long long s_geo;   // geometric survivors of the other wavefronts (several per replica)

    for (;;) {
        // Claim a work item: a class member that no block holds and that has chunks left.  Wave
        // 0 probes the members 64 at a time from a rotating cursor and takes one with a
        // compare-and-swap; when none is free the block ends (the members left are held by
        // running blocks, which take them on).  Every probe is bounded, so the grid drains.
        if (wave == 0) {
            int got = -1, chunk = 0;
            const uint32_t start = (uint32_t)__shfl(lane == 0 ? atomicAdd(A.queue, 1) : 0, 0);
            for (int base = 0; base < A.count && got < 0; base += 64) {
                const int m = (int)((start + (uint32_t)(base + lane)) % (uint32_t)A.count);
                const bool in = base + lane < A.count;
                const int st = in ? __hip_atomic_load(A.state + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 1;
                uint64_t free_mask = __ballot(in && !(st & 1) && (st >> 1) < A.nchunks);
                while (free_mask && got < 0) {
                    const int l = __ffsll((unsigned long long)free_mask) - 1;
                    int won = 0;
                    if (lane == l) {
                        int expect = st;
                        won = __hip_atomic_compare_exchange_strong(A.state + m, &expect, st | 1, __ATOMIC_RELAXED,
                                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    if (__shfl(won, l)) {
                        got = __shfl(m, l);
                        chunk = __shfl(st >> 1, l);
                    }
                    free_mask &= free_mask - 1;
                }
            }
            if (lane == 0) {
                s_item = got;
                s_chunk = chunk;
                // acquire: the previous holder's stores (any XCD) are visible from here on
                if (got >= 0) __threadfence();
            }
        }
        bsync();
        const int member = s_item;
        if (member < 0) break;
        const int r = A.order[member];
        if (tid == 0) {
            R = A.rep[r];
            s_geo = 0;
        }
        bsync();
        const int chunk = s_chunk;
        const int legs_now = min(A.chunk_legs, A.legs - chunk * A.chunk_legs);
        const int n = R.n, cps = R.cps;
        const double L = R.L, side = R.side;
        const double *P = A.pos64 + R.off;
        const double inv_step = 0x1p16 / L;
        for (int i = tid; i < n; i += NTH) {
            qx[i] = quant1(P[i], inv_step);
            qy[i] = quant1(P[(size_t)n + i], inv_step);
            qz[i] = quant1(P[2 * (size_t)n + i], inv_step);
        }
        const int ncell = cps ? cps * cps * cps : 0;
        // counting sort of the units by cell; the counts and cursors live in the contender
        // area, which is idle here (4 * 512 bytes fit in it)
        for (int c = tid; c < ncell; c += NTH) ccount[c] = 0;
        if (tid < 3) s_ap[tid] = P[(size_t)tid * n + R.active];
        if (cps && tid <= cps) s_bounds[tid] = A.bounds[R.bound_off + tid];
        bsync();
        if (cps) {
            auto unit_cell = [&](int i) {
                return bcell1(P[i], R.side, cps) +
                       cps * (bcell1(P[(size_t)n + i], R.side, cps) + cps * bcell1(P[2 * (size_t)n + i], R.side, cps));
            };
            for (int i = tid; i < n; i += NTH)
                if (i != R.active) atomicAdd(&ccount[unit_cell(i)], 1);
            bsync();
            if (wave == 0) {
                // exclusive scan over the cells: each lane a run of up to 8, then across lanes
                const int per = (ncell + 63) / 64;
                int sum = 0;
                for (int k = 0; k < per; k++) {
                    const int c = lane * per + k;
                    if (c < ncell) sum += ccount[c];
                }
                int incl = sum;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int v = __shfl_up(incl, o, 64);
                    if (lane >= o) incl += v;
                }
                int run = incl - sum;
                for (int k = 0; k < per; k++) {
                    const int c = lane * per + k;
                    if (c < ncell) {
                        const int t = ccount[c];
                        cs[c] = (int16_t)run;
                        ccount[c] = run;
                        run += t;
                    }
                }
                if (lane == 63) cs[ncell] = (int16_t)incl;
            }
            bsync();
            for (int i = tid; i < n; i += NTH)
                if (i != R.active) cu[atomicAdd(&ccount[unit_cell(i)], 1)] = (int16_t)i;
        }
        const float Lf = (float)L;
        const float step = (float)(L * 0x1p-16);
        const float E = Lf * 0x1p-15f;
        const double rc2 = R.cutoff * R.cutoff;
        const float rc2_hi = (float)rc2 * (1.f + 0x1p-22f);
        const int nf = A.fac.n_factors;
        // per-factor constants (the second factor's equal the first's when there is one factor)
        const double pw0 = A.fac.power[0], pw1 = nf > 1 ? A.fac.power[1] : pw0;
        const double pk0 = A.fac.prefactor[0], pk1 = nf > 1 ? A.fac.prefactor[1] : pk0;
        const double urc0 = pk0 / cr_pow(rc2, pw0 / 2.0), urc1 = pk1 / cr_pow(rc2, pw1 / 2.0);
        const float vrc0 = (float)fabs(urc0) * (1.f - 0x1p-20f), vrc1 = (float)fabs(urc1) * (1.f - 0x1p-20f);
        const float kf0 = (float)pk0, kf1 = (float)pk1;
        const int pf0 = (int)pw0, pf1 = (int)pw1;
        const float inv_beta_lo = (float)(1.0 / A.beta) * (1.f - 0x1p-22f);
        // pre-filter constants (cand_near): (rc + 2E)^2 rounded up, half the box
        const float RB = ((float)R.cutoff * (1.f + 0x1p-22f) + 2.f * E) * (1.f + 0x1p-20f);
        const float RB2 = RB * RB * (1.f + 0x1p-20f);
        const float halfL = 0.5f * Lf, E2 = 2.f * E;
        // A separation within 2E of +-L/2 on some axis has |s_q| >= L/2 - 2E.  When that exceeds
        // the capsule's reach rc + 2E ahead of the largest horizon (the chain length, rounded
        // up as hf is), such a candidate fails the capsule test on either image, so neither
        // pass needs the ambiguity test.
        const float hf_max = (float)A.chain_length * (1.f + 0x1p-22f) + 0x1p-60f;
        const bool check_amb = !(halfL - E2 > (hf_max + RB) * (1.f + 0x1p-20f));
        // with 3 cells per side the 27 nearby cells are the whole box: candidate idx is unit idx
        // or idx + 1, past the active unit
        const bool whole_box = cps == 0 || cps == 3;
        // Replica state through the chunk, in registers (with one wavefront per replica the LDS
        // copy R is only read at the start and written at the end; with several, wavefront 0
        // publishes it to LDS after each commit and every wavefront reloads it)
        int st_a = R.active, st_d = R.direction, st_err = R.err;
        uint64_t st_leg = R.leg;
        double st_cl = R.chain_left;
        double st_ap[3] = {s_ap[0], s_ap[1], s_ap[2]};
        double st_h = 0.0, st_cb = 0.0;
        int st_M = 0, st_pc[3] = {0, 0, 0};
        // Nearby cells that cannot meet the capsule of the next leg are left out of its candidate
        // list (their count is taken as zero): the box distance from the segment [0, h] (h
        // widened to cover hf) must be at least rc + 2E rounded up, plus 2E for the position
        // slack of the fixed-point test and L 2^-18 for the FP32 box bounds (each rounded once
        // from an FP64 difference below L, then a few FP32 operations).  Not in the all-FP64
        // mode (which inverts every candidate of the 27 cells) nor where the ambiguity tests
        // are on; the audit mode keeps it and checks every candidate of the 27 cells, culled
        // or not.  (The radius is recomputed in each call from the cutoff in LDS, so that it
        // does not hold a register through the leg.)
        const bool cull = A.mode != 1 && !check_amb;
        // the estimators, summed leg by leg in the oracle's order (wavefront 0)
        double dest[6] = {R.est[0], R.est[1], R.est[2], R.est[3], R.est[4], R.est[5]};
        unsigned long long dst[4] = {0ull, 0ull, 0ull, 0ull};  // statistics increments (wavefront 0)
        // The next leg's 27 nearby cells with an inclusive scan of their counts (not needed when
        // they are the whole box), its cell, horizon and candidate count, and cleared contender
        // counters: wavefront 0, before the first leg and right after each commit (so that no
        // separate barrier precedes pass 1)
        auto prepare_leg = [&](const double *pa, int dn_, double cl) {
            int pc[3] = {0, 0, 0};
            if (cps) {
#pragma unroll
                for (int q = 0; q < 3; q++) pc[q] = bcell1(pa[q], side, cps);
            }
            double hh = cl, cb = INFINITY;
            if (cps) {
                cb = s_bounds[pc[dn_] + 1] - pa[dn_];
                if (cb < hh) hh = cb;
            }
            int Mv = n - 1, Mf = n - 1;
            if (cps && !whole_box) {
                int cnt = 0, cntf = 0;
                if (lane < 27) {
                    const int off[3] = {lane % 3 - 1, (lane / 3) % 3 - 1, lane / 9 - 1};
                    int cc[3];
#pragma unroll
                    for (int q = 0; q < 3; q++) cc[q] = (pc[q] + off[q] + cps) % cps;
                    const int c = cc[0] + cps * (cc[1] + cps * cc[2]);
                    s_nearby[lane] = cs[c];
                    cntf = cs[c + 1] - cs[c];
                    bool live = true;
                    if (cull) {
                        const float rb = ((float)R.cutoff * (1.f + 0x1p-22f) + 2.f * E) * (1.f + 0x1p-20f);
                        const float cull_r = (rb * (1.f + 0x1p-20f) + 2.f * E + Lf * 0x1p-18f) * (1.f + 0x1p-20f);
                        const float cull_r2 = cull_r * cull_r * (1.f + 0x1p-20f);
                        // the cell's box relative to the active unit (shifted by L where the
                        // neighbour wraps), against the point 0 across the motion and the
                        // segment [0, h] along it
                        const float top = (float)hh * (1.f + 0x1p-20f) + 0x1p-50f;
                        float g2 = 0.f;
#pragma unroll
                        for (int q = 0; q < 3; q++) {
                            const double shift = (off[q] < 0 && pc[q] == 0) ? -L : ((off[q] > 0 && pc[q] == cps - 1) ? L : 0.0);
                            const float lo = (float)(s_bounds[cc[q]] - pa[q] + shift);
                            const float hi = (float)(s_bounds[cc[q] + 1] - pa[q] + shift);
                            const float gap = fmaxf(fmaxf(lo - (q == dn_ ? top : 0.f), -hi), 0.f);
                            g2 += gap * gap;
                        }
                        live = g2 < cull_r2;
                    }
                    cnt = live ? cntf : 0;
                }
                // inclusive scans of the 27 counts (kept and all): within each 16-lane row by DPP
                // row shifts (a lane whose source is outside the row adds 0), then lanes 16-26
                // add row 0's sum
                int incl = cnt, inclf = cntf;
                incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xF, 0xF, false);
                inclf += __builtin_amdgcn_update_dpp(0, inclf, 0x111, 0xF, 0xF, false);
                incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xF, 0xF, false);
                inclf += __builtin_amdgcn_update_dpp(0, inclf, 0x112, 0xF, 0xF, false);
                incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xF, 0xF, false);
                inclf += __builtin_amdgcn_update_dpp(0, inclf, 0x114, 0xF, 0xF, false);
                incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xF, 0xF, false);
                inclf += __builtin_amdgcn_update_dpp(0, inclf, 0x118, 0xF, 0xF, false);
                const int row0 = __builtin_amdgcn_readlane(incl, 15), row0f = __builtin_amdgcn_readlane(inclf, 15);
                if (lane >= 16) {
                    incl += row0;
                    inclf += row0f;
                }
                if (lane < 27) s_pref[lane + 1] = incl;
                if (lane == 0) s_pref[0] = 0;
                if (A.mode == 2) {
                    if (lane < 27) s_preff[lane + 1] = inclf;
                    if (lane == 0) s_preff[0] = 0;
                }
                Mv = __builtin_amdgcn_readlane(incl, 26);
                Mf = __builtin_amdgcn_readlane(inclf, 26);
            }
            st_h = hh;
            st_cb = cb;
            st_M = Mv;
#pragma unroll
            for (int q = 0; q < 3; q++) st_pc[q] = pc[q];
            if (lane == 0) {
                s_nc = 0;
                s_overflow = 0;
                s_Mfull = Mf;   // the statistics count every candidate of the 27 cells
                if (WPR > 1) {
                    s_h = hh;
                    s_cb = cb;
                    s_M = Mv;
                }
            }
        };
        if (wave == 0) prepare_leg(st_ap, st_d, st_cl);
        bsync();
        if (WPR > 1) {
            st_h = s_h;
            st_cb = s_cb;
            st_M = s_M;
        }
        for (int leg_i = 0; leg_i < legs_now && !st_err; leg_i++) {
            const int a = st_a, d = st_d;
            const uint64_t leg = st_leg;
            const double ap[3] = {st_ap[0], st_ap[1], st_ap[2]};
            const int M = st_M;
            const double h = st_h;
            const uint32_t qa0 = quant1(ap[0], inv_step), qa1 = quant1(ap[1], inv_step), qa2 = quant1(ap[2], inv_step);
            const float hf = (float)h * (1.f + 0x1p-22f) + 0x1p-60f;
            BKey best = {INFINITY, 3, 0x7ffffff, 0};   // j < 2^27: the reduction packs (rank, j, f)
            double bestpos[3] = {0.0, 0.0, 0.0};
            unsigned geo_pass = 0;   // geometric survivors of this wavefront (uniform)
            if (A.mode != 1) {
                // ---- pass 1b: bounds of the queued candidates of this wavefront
                int nc = 0;   // contenders so far (one wavefront per replica: uniform register)
                auto bound_queue = [&](int cnt) {
                    const bool act = lane < cnt;
                    int j = 0;
                    uint32_t w[4] = {0u, 0u, 0u, 0u};
                    bool con[2] = {false, false};
                    if (act) {
                        j = queue[lane];
                        Geo g;
                        const bool amb = cand_geometry(qsep(qx[j], qa0, step), qsep(qy[j], qa1, step), qsep(qz[j], qa2, step),
                                                       d, Lf, E, check_amb, g);
                        philox4x32_10((uint32_t)leg, (uint32_t)(leg >> 32), (uint32_t)j, 0u, (uint32_t)r, A.seed_lo, w);
#pragma unroll
                        for (int f = 0; f < 2; f++) {
                            if (f >= nf) break;
                            float t = 0.f;
                            if (!amb) {
                                if (g.perp2_lo >= rc2_hi) continue;
                                if (factor_geometric_floor(g, pick(f, kf0, kf1), rc2_hi) > hf) continue;
                                t = factor_lower_bound(g, pick(f, kf0, kf1), pick(f, pf0, pf1), rc2_hi, pick(f, vrc0, vrc1),
                                                       dU_lower(pick(f, w[0], w[2]), inv_beta_lo));
                            }
                            con[f] = t <= hf;
                        }
                    }
                    // contender slots: with one wavefront per replica from ballots over the whole
                    // (converged) wavefront, else from the workgroup's LDS counter
#pragma unroll
                    for (int f = 0; f < 2; f++) {
                        int slot;
                        if (WPR == 1) {
                            const unsigned long long m = __ballot(con[f]);
                            slot = nc + __popcll(m & ((1ull << lane) - 1ull));
                            nc += __popcll(m);
                            if (!con[f]) continue;
                        } else {
                            if (!con[f]) continue;
                            slot = atomicAdd(&s_nc, 1);
                        }
                        if (slot >= B_CONTENDERS) {
                            s_overflow = 1;
                            continue;
                        }
                        cont_id[slot] = (j << 1) | f;
                        cont_w[2 * slot] = pick(f, w[0], w[2]);
                        cont_w[2 * slot + 1] = pick(f, w[1], w[3]);
                    }
                };
                // ---- pass 1a: every candidate against the capsule, survivors queued; three
                // chunks of 64 per wavefront and iteration, so that their LDS reads (cell search,
                // unit, position) overlap
                int qn = 0;
                auto append = [&](bool surv, int j) {
                    const unsigned long long m = __ballot(surv);
                    const int before = __popcll(m & ((1ull << lane) - 1ull));
                    if (surv) queue[qn + before] = j;
                    qn += __popcll(m);
                    if (qn >= 64) {
                        bound_queue(64);
                        const int rem = qn - 64;
                        const int v = lane < rem ? queue[64 + lane] : 0;
                        if (lane < rem) queue[lane] = v;
                        qn = rem;
                    }
                    geo_pass += (unsigned)__popcll(m);
                };
                for (int base = wave * 64; base < M; base += 3 * NTH) {
                    const int i0 = base + lane, i1 = base + NTH + lane, i2 = base + 2 * NTH + lane;
                    int j0, j1, j2;
                    if (whole_box) {
                        j0 = i0 < a ? i0 : i0 + 1;
                        j1 = i1 < a ? i1 : i1 + 1;
                        j2 = i2 < a ? i2 : i2 + 1;
                    } else {
                        candidate_units3(min(i0, M - 1), min(i1, M - 1), min(i2, M - 1), s_pref, s_nearby, cu, j0, j1, j2);
                    }
                    bool s0 = false, s1 = false, s2 = false;
                    if (i0 < M)
                        s0 = cand_near(qsep(qx[j0], qa0, step), qsep(qy[j0], qa1, step), qsep(qz[j0], qa2, step), d, halfL,
                                       E2, RB2, hf, check_amb);
                    if (i1 < M)
                        s1 = cand_near(qsep(qx[j1], qa0, step), qsep(qy[j1], qa1, step), qsep(qz[j1], qa2, step), d, halfL,
                                       E2, RB2, hf, check_amb);
                    if (i2 < M)
                        s2 = cand_near(qsep(qx[j2], qa0, step), qsep(qy[j2], qa1, step), qsep(qz[j2], qa2, step), d, halfL,
                                       E2, RB2, hf, check_amb);
                    append(s0, j0);
                    append(s1, j1);
                    append(s2, j2);
                }
                if (qn > 0) bound_queue(qn);
                if (WPR == 1 && lane == 0) s_nc = nc;
                bsync();
            }
            const bool all_exact = (A.mode == 1) || s_overflow;
            // ---- pass 2: exact FP64 inversion of the contenders (or of every candidate)
            const int nwork = all_exact ? M * nf : min(s_nc, B_CONTENDERS);
            for (int w = tid; w < nwork; w += NTH) {
                int j, f;
                double pj[3];
                uint32_t wa, wb;
                if (all_exact) {
                    f = w % nf;
                    const int idx = w / nf;
                    j = whole_box ? (idx < a ? idx : idx + 1) : candidate_unit(idx, cps, a, s_pref, s_nearby, cu);
                    pj[0] = P[j]; pj[1] = P[(size_t)n + j]; pj[2] = P[2 * (size_t)n + j];
                    uint32_t ww[4];
                    philox4x32_10((uint32_t)leg, (uint32_t)(leg >> 32), (uint32_t)j, 0u, (uint32_t)r, A.seed_lo, ww);
                    wa = pick(f, ww[0], ww[2]);
                    wb = pick(f, ww[1], ww[3]);
                } else {
                    j = cont_id[w] >> 1;
                    f = cont_id[w] & 1;
                    // the FP64 position from global memory (its latency overlaps the log of
                    // the draw inside exact_event)
                    pj[0] = P[j]; pj[1] = P[(size_t)n + j]; pj[2] = P[2 * (size_t)n + j];
                    wa = cont_w[2 * w];
                    wb = cont_w[2 * w + 1];
                }
                const double dx = exact_event(wa, wb, A.beta, bmin_image(pj[0] - ap[0], L), bmin_image(pj[1] - ap[1], L),
                                              bmin_image(pj[2] - ap[2], L), d, pick(f, pf0, pf1) == 12, pick(f, pk0, pk1),
                                              rc2, pick(f, urc0, urc1));
                BKey c = {dx, 0, j, f};
                if (bkey_lt(c, best)) {
                    best = c;
                    bestpos[0] = pj[0]; bestpos[1] = pj[1]; bestpos[2] = pj[2];
                }
            }
            if (A.mode == 2) {
                // audit: every candidate of the 27 cells (or of the whole box), including those in
                // cells culled from the leg's list, whose exact displacement is <= h must be a
                // contender (before the selection's barrier: after it, wavefront 0 already
                // prepares the next leg)
                const int Ma = whole_box ? M : s_Mfull;
                for (int idx = tid; idx < Ma * nf; idx += NTH) {
                    const int f = idx % nf;
                    const int ci = idx / nf;
                    const int j = whole_box ? (ci < a ? ci : ci + 1) : candidate_unit(ci, cps, a, s_preff, s_nearby, cu);
                    uint32_t ww[4];
                    philox4x32_10((uint32_t)leg, (uint32_t)(leg >> 32), (uint32_t)j, 0u, (uint32_t)r, A.seed_lo, ww);
                    const double dx = exact_event(pick(f, ww[0], ww[2]), pick(f, ww[1], ww[3]), A.beta,
                                                  bmin_image(P[j] - ap[0], L), bmin_image(P[(size_t)n + j] - ap[1], L),
                                                  bmin_image(P[2 * (size_t)n + j] - ap[2], L), d, pick(f, pf0, pf1) == 12,
                                                  pick(f, pk0, pk1), rc2, pick(f, urc0, urc1));
                    if (dx <= h) {
                        bool found = false;
                        for (int q = 0; q < min(s_nc, B_CONTENDERS); q++)
                            if (cont_id[q] == ((j << 1) | f)) found = true;
                        if (!found && !s_overflow) atomicAdd(A.audit, 1ull);
                    }
                }
            }
            BKey mine = best;
            if (tid == 0) {
                BKey c1k = {st_cb, 1, 0, 0};
                if (bkey_lt(c1k, best)) best = c1k;
                BKey c2k = {st_cl, 2, 0, 0};
                if (bkey_lt(c2k, best)) best = c2k;
            }
            const int wbest = block_min<WPR>(best, mine, bestpos, sk, s_newpos, nwork);
            if (wave == 0) {
                dst[0] += (unsigned long long)s_Mfull;   // every candidate of the 27 cells
                dst[2] += (unsigned long long)min(s_nc, B_CONTENDERS);
                dst[3] += (unsigned long long)nwork;
            }
            if (WPR == 1) dst[1] += geo_pass;
            else if (lane == 0 && geo_pass) atomicAdd(&s_geo, (unsigned long long)geo_pass);
            // ---- commit (wavefront 0; the winner's position is in the winning lane's registers
            // or, with several wavefronts, in s_newpos): the values below are uniform across the
            // wavefront, a ballot finds the new active unit's slot in its cell list, lane 0 writes
            if (wave == 0) {
                double *Pw = A.pos64 + R.off;
                double x;
                if (best.rank == 1) {
                    int c = st_pc[d];
                    x = c + 1 < cps ? s_bounds[c + 1] : 0.0;
                } else {
                    x = ap[d] + best.d;
                    if (x >= L) x -= L;
                }
                int newa = a;
                double np[3] = {ap[0], ap[1], ap[2]};
                np[d] = x;
                const double pold[3] = {np[0], np[1], np[2]};
                if (best.rank == 0) {
                    newa = best.j;
                    if (WPR == 1) {
#pragma unroll
                        for (int k = 0; k < 3; k++)
                            np[k] = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(bestpos[k]), wbest),
                                                     __builtin_amdgcn_readlane(__double2loint(bestpos[k]), wbest));
                    } else {
                        np[0] = s_newpos[3 * wbest]; np[1] = s_newpos[3 * wbest + 1]; np[2] = s_newpos[3 * wbest + 2];
                    }
                } else if (best.rank == 2) {
                    uint32_t w[4];
                    philox4x32_10((uint32_t)leg, (uint32_t)(leg >> 32), 0xFFFFFFFFu, 1u, (uint32_t)r, A.seed_lo, w);
                    newa = (int)(((uint64_t)w[0] * (uint64_t)n) >> 32);
                    if (newa != a) {
                        np[0] = Pw[newa]; np[1] = Pw[(size_t)n + newa]; np[2] = Pw[2 * (size_t)n + newa];
                    }
                }
                int ca = 0, cn = 0, q = -1;
                if (newa != a && cps) {
                    ca = bcell1(pold[0], side, cps) + cps * (bcell1(pold[1], side, cps) + cps * bcell1(pold[2], side, cps));
                    cn = bcell1(np[0], side, cps) + cps * (bcell1(np[1], side, cps) + cps * bcell1(np[2], side, cps));
                    // the new active unit's slot in its cell's segment
                    const int b1 = cs[cn + 1];
                    for (int base = cs[cn]; base < b1 && q < 0; base += 64) {
                        const unsigned long long mask = __ballot(base + lane < b1 && cu[base + lane] == newa);
                        if (mask) q = base + __ffsll((long long)mask) - 1;
                    }
                }
                // the replica's state, alike in every lane of wavefront 0
                dest[1] += best.d;
                if (best.rank == 0) {
                    dest[0] += bmin_image(np[d] - ap[d], L) - best.d;
                    dest[2] += 1.0;
                    dest[4] += 1.0;
                    st_cl -= best.d;
                } else if (best.rank == 1) {
                    dest[3] += 1.0;
                    st_cl -= best.d;
                } else {
                    dest[5] += 1.0;
                    st_cl = A.chain_length;
                    st_d = (d + 1) % 3;
                }
                if (newa != a && cps && q < 0) st_err = 2;
                st_a = newa;
                st_leg = leg + 1;
#pragma unroll
                for (int k = 0; k < 3; k++) st_ap[k] = np[k];
                if (lane == 0) {
                    Pw[(size_t)d * n + a] = x;
                    qx[a] = quant1(pold[0], inv_step);
                    qy[a] = quant1(pold[1], inv_step);
                    qz[a] = quant1(pold[2], inv_step);
                    // the old active unit takes the new one's slot (and is moved to the cell it
                    // ended in below)
                    if (newa != a && cps && q >= 0) cu[q] = (int16_t)a;
                    if (A.log && leg < (uint64_t)A.log_legs) {
                        ecmc_batch_event ev;
                        ev.kind = best.rank;
                        ev.active = a;
                        ev.target = best.rank == 0 ? best.j : (best.rank == 2 ? newa : -1);
                        ev.factor = best.rank == 0 ? best.f : 0;
                        ev.displacement = best.d;
                        ev.x = x;
                        A.log[(size_t)r * A.log_legs + leg] = ev;
                    }
                    if (WPR > 1) {
                        R.active = st_a;
                        R.direction = st_d;
                        R.leg = st_leg;
                        R.chain_left = st_cl;
                        R.err = st_err;
                        s_ap[0] = np[0]; s_ap[1] = np[1]; s_ap[2] = np[2];
                    }
                }
                // lane 0's LDS writes above precede the other lanes' reads below (one wavefront,
                // LDS operations in order; the fence keeps the compiler from reordering them)
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (newa != a && cps && q >= 0 && ca != cn) {
                    csr_move(cu, cs, q, cn, ca, a, lane);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                }
                prepare_leg(st_ap, st_d, st_cl);
            }
            bsync();
            if (WPR > 1) {
                st_a = R.active;
                st_d = R.direction;
                st_leg = R.leg;
                st_cl = R.chain_left;
                st_err = R.err;
                st_ap[0] = s_ap[0]; st_ap[1] = s_ap[1]; st_ap[2] = s_ap[2];
                st_h = s_h;
                st_cb = s_cb;
                st_M = s_M;
            }
        }
        if (tid == 0) {
            R.active = st_a;
            R.direction = st_d;
            R.leg = st_leg;
            R.chain_left = st_cl;
            R.err = st_err;
            for (int q = 0; q < 6; q++) R.est[q] = dest[q];
            R.stats[0] += (int64_t)dst[0];
            R.stats[1] += (int64_t)(dst[1] + s_geo);
            R.stats[2] += (int64_t)dst[2];
            R.stats[3] += (int64_t)dst[3];
            A.rep[r] = R;
        }
        bsync();   // every store of this chunk (positions, log, state) has completed
        if (tid == 0) {
            // release: make this block's stores visible device-wide, then free the member with
            // one more chunk done
            // (an exchange, whose return is awaited, so this block's next probe sees the update)
            __threadfence();
            const int prev = __hip_atomic_exchange(A.state + member, (chunk + 1) << 1, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
            if (prev != ((chunk << 1) | 1)) A.rep[r].err = 3;   // the claim protocol was broken
        }
    }
}

}  // namespace ecmc
