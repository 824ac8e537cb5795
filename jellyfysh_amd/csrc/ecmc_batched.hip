// ecmc_batched.hip -- thousands of independent ECMC replicas of uneven size advanced by one
// persistent kernel per size class.
//
// Each workgroup (WPR wavefronts) pulls replicas from a queue sorted largest first (LPT)
// and runs `legs` legs on each.  A replica's FP32 positions and its cell lists live in LDS;
// its FP64 positions stay in HBM, where only the active unit's row changes.  One leg:
//   1. horizon h = min(displacement to the next cell wall, rest of the chain), FP64;
//   2. every candidate (occupants of the 27 nearby cells, or all units when the box holds
//      fewer than three cells per side) and factor gets an FP32 lower bound on its event
//      displacement, with a proven downward margin; only those whose bound is <= h enter
//      a contender list in LDS;
//   3. the contenders are inverted exactly in FP64 (correctly rounded pow/log) and a
//      workgroup min-reduce over (displacement, kind, target, factor) picks the event;
//   4. thread 0 commits it (lifting, wall snap, or end of chain with a Philox-drawn new
//      active unit) and updates the cell lists.
// Because every bound is <= the exact FP64 value, the committed event is the one a full
// FP64 evaluation of every candidate selects; the audit mode checks that on the device.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <vector>

#include "../../include/jellyfysh_amd.h"
#include "ecmc_common.h"
#include "ecmc_math.h"
#include "ecmc_potentials.h"

namespace ecmc {

constexpr int B_MAX_WPR = 8;
constexpr int B_CELL_CAP = 64;       // occupants per cell
constexpr int B_CONTENDERS = 256;    // contender list capacity per leg
constexpr int B_MAX_CPS = 8;         // cells per side handled in LDS (512 cells)

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
    int32_t count;               // replicas in this class
    int32_t nmax, cellmax;       // LDS sizing of this class
    int32_t legs;
    int32_t mode;                // 0 filtered, 1 all candidates exact, 2 filtered + audit
    int32_t log_legs;            // legs logged per replica (0 = none)
    uint32_t seed_lo, pad;
    double beta, chain_length;
    FactorDev fac;
    int32_t *queue;              // work counter of this class
    ecmc_batch_event *log;       // n_replicas x log_legs
    unsigned long long *audit;   // bound violations
};

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) --------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
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
// Every quantity is bracketed [lo, hi] and each FP32 operation is covered by a relative
// margin (ETA for products/quotients/roots, ETA_POW for the exp2/log2 power), far above
// FP32 rounding (2^-24 per operation) and the documented accuracy of log2f/exp2f.
constexpr float ETA = 0x1p-17f, ETA_POW = 0x1p-14f;
__device__ __forceinline__ float up(float x) { return x >= 0.f ? x * (1.f + ETA) : x * (1.f - ETA); }
__device__ __forceinline__ float dn(float x) { return x >= 0.f ? x * (1.f - ETA) : x * (1.f + ETA); }
__device__ __forceinline__ float sq_lo(float a_abs_lo) { return dn(a_abs_lo * a_abs_lo); }
__device__ __forceinline__ float sq_hi(float a_abs_hi) { return up(a_abs_hi * a_abs_hi); }

// V(r2) = K / r2^(p/2) with p in {6, 12} (exact integer powers by multiplication); returns
// an upper (hi=true) or lower bound for r2 given as an upper/lower bound correspondingly.
__device__ __forceinline__ float vpow_bound(float K, float r2, int half_p, bool hi) {
    float r3 = r2 * r2 * r2;
    float rp = half_p == 6 ? r3 * r3 : r3;
    float v = K / rp;
    // five roundings at most
    return hi ? v * (1.f + 8.f * 0x1p-24f) * (1.f + ETA) : v * (1.f - 8.f * 0x1p-24f) * (1.f - ETA);
}

// (K / V)^(2/p) bounded up (hi) or down.
__device__ __forceinline__ float rstar2_bound(float K, float V, float two_over_p, bool hi) {
    float q = K / V;
    float r = exp2f(two_over_p * log2f(q));
    return hi ? r * (1.f + ETA_POW) : r * (1.f - ETA_POW);
}

struct Geo {
    float sd_lo, sd_hi, perp2_lo, perp2_hi, r2_lo, r2_hi;
};

// Lower bound of the exact displacement of one factor, or +INF when the factor surely has
// no event, given the exponential draw's lower bound dU_lo.  `k` is the prefactor (charges
// are one), `p` the power (6 or 12), rc2 the cutoff squared (double, exact reference).
__device__ float factor_lower_bound(const Geo &g, float k, int p, float rc2_lo, float rc2_hi, float vrc_lo,
                                    float vrc_hi, float dU_lo) {
    const float INF = __int_as_float(0x7f800000);
    const int half_p = p / 2;
    const float two_over_p = 2.f / (float)p;
    if (g.perp2_lo >= rc2_hi) return INF;
    if (k > 0.f) {
        if (g.sd_hi <= 0.f) return INF;
        // U_eff = max(U(r2), U(rc)), U_max = U(perp2)
        float ueff_lo = fmaxf(vpow_bound(k, g.r2_hi, half_p, false), vrc_lo);
        float umax_hi = vpow_bound(k, fmaxf(g.perp2_lo, 1e-30f), half_p, true);
        float room_hi = up(umax_hi - ueff_lo);
        if (dU_lo >= room_hi) return INF;
        float ut_lo = dn(ueff_lo + dU_lo);
        float rs2_hi = rstar2_bound(k, ut_lo, two_over_p, true);
        float x_hi = up(sqrtf(fmaxf(up(rs2_hi - g.perp2_lo), 0.f)));
        float d = dn(g.sd_lo) - x_hi;
        d = d - fabsf(d) * 0x1p-22f;
        return fmaxf(d, 0.f);
    }
    // attractive: V = |k| / r^p, the event comes while receding, after the closest approach
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
    float y_lo = dn(sqrtf(fmaxf(dn(rs2_lo - g.perp2_hi), 0.f)));
    float d = g.sd_lo + y_lo;
    d = d - fabsf(d) * 0x1p-22f - 0x1p-40f;
    return fmaxf(d, 0.f);
}

// Smallest displacement any event of the factor can have, from geometry alone (no draw):
// a repulsive factor first has to reach the cutoff sphere, an attractive one first passes
// the closest approach.
__device__ __forceinline__ float factor_geometric_floor(const Geo &g, float k, float rc2_hi) {
    if (k > 0.f) {
        if (g.r2_lo < rc2_hi) return 0.f;
        float entry_hi = up(sqrtf(fmaxf(up(rc2_hi - g.perp2_lo), 0.f)));
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

template <int WPR>
__device__ void block_min(BKey &b, BKey *sk) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        BKey o;
        o.d = __shfl_xor(b.d, off, 64);
        o.rank = __shfl_xor(b.rank, off, 64);
        o.j = __shfl_xor(b.j, off, 64);
        o.f = __shfl_xor(b.f, off, 64);
        if (bkey_lt(o, b)) b = o;
    }
    if (WPR > 1) {
        if (lane == 0) sk[wave] = b;
        __syncthreads();
        b = sk[0];
#pragma unroll
        for (int w = 1; w < WPR; w++)
            if (bkey_lt(sk[w], b)) b = sk[w];
        __syncthreads();
    }
}

__device__ __forceinline__ void bsync() { __syncthreads(); }

// exact FP64 displacement of factor f of target j
__device__ double exact_displacement(const BatchArgs &A, const ReplicaDev &R, const double *ap, int d, int j, int f,
                                     uint64_t leg, uint32_t replica) {
    double s[3];
    const double *pj = A.pos64 + R.off;
#pragma unroll
    for (int q = 0; q < 3; q++) s[q] = bmin_image(pj[(size_t)q * R.n + j] - ap[q], R.L);
    uint32_t w[4];
    philox4x32_10((uint32_t)leg, (uint32_t)(leg >> 32), (uint32_t)j, 0u, replica, A.seed_lo, w);
    double u = unit53(w[2 * f], w[2 * f + 1]);
    double dU = -cr_log(1.0 - u) / A.beta;
    return ipt_displacement<3>(d, 1.0, s, 1.0, 1.0, A.fac.power[f], A.fac.prefactor[f], R.cutoff, dU);
}

template <int WPR>
__global__ __launch_bounds__(64 * WPR) void batched_kernel(BatchArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NTH = 64 * WPR;
    const int tid = threadIdx.x;
    float *px = (float *)smem;
    float *py = px + A.nmax;
    float *pz = py + A.nmax;
    int32_t *cell_n = (int32_t *)(pz + A.nmax);
    int16_t *cell_l = (int16_t *)(cell_n + A.cellmax);
    int32_t *cont_id = (int32_t *)(cell_l + (size_t)A.cellmax * B_CELL_CAP + (((size_t)A.cellmax * B_CELL_CAP) & 1));
    __shared__ BKey sk[B_MAX_WPR];
    __shared__ int s_item, s_nc, s_overflow, s_nearby[27], s_pref[28], s_M;
    __shared__ double s_ap[3], s_h, s_cb;
    __shared__ ReplicaDev R;
    __shared__ unsigned long long s_stats[4];

    for (;;) {
        if (tid == 0) s_item = atomicAdd(A.queue, 1);
        bsync();
        const int item = s_item;
        if (item >= A.count) break;
        const int r = A.order[item];
        if (tid == 0) {
            R = A.rep[r];
            for (int q = 0; q < 4; q++) s_stats[q] = 0;
        }
        bsync();
        const int n = R.n, cps = R.cps;
        const double L = R.L;
        const double *P = A.pos64 + R.off;
        for (int i = tid; i < n; i += NTH) {
            px[i] = (float)P[i];
            py[i] = (float)P[(size_t)n + i];
            pz[i] = (float)P[2 * (size_t)n + i];
        }
        const int ncell = cps ? cps * cps * cps : 0;
        for (int c = tid; c < ncell; c += NTH) cell_n[c] = 0;
        bsync();
        if (cps) {
            for (int i = tid; i < n; i += NTH) {
                if (i == R.active) continue;
                int c = bcell1(P[i], R.side, cps) + cps * (bcell1(P[(size_t)n + i], R.side, cps) +
                                                           cps * bcell1(P[2 * (size_t)n + i], R.side, cps));
                int slot = atomicAdd(&cell_n[c], 1);
                if (slot < B_CELL_CAP) cell_l[(size_t)c * B_CELL_CAP + slot] = (int16_t)i;
            }
            bsync();
            if (tid == 0)
                for (int c = 0; c < ncell; c++)
                    if (cell_n[c] > B_CELL_CAP) R.err = 1;
        }
        const float Lf = (float)L;
        const float E = Lf * 0x1p-21f;
        const double rc2 = R.cutoff * R.cutoff;
        const float rc2_lo = (float)rc2 * (1.f - 0x1p-22f), rc2_hi = (float)rc2 * (1.f + 0x1p-22f);
        float vrc_lo[2], vrc_hi[2];
        for (int f = 0; f < A.fac.n_factors; f++) {
            double v = fabs(A.fac.prefactor[f]) / cr_pow(rc2, A.fac.power[f] / 2.0);
            vrc_lo[f] = (float)v * (1.f - 0x1p-20f);
            vrc_hi[f] = (float)v * (1.f + 0x1p-20f);
        }
        bsync();
        for (int leg_i = 0; leg_i < A.legs && !R.err; leg_i++) {
            const int a = R.active, d = R.direction;
            const uint64_t leg = R.leg;
            if (tid == 0) {
                s_ap[0] = P[a];
                s_ap[1] = P[(size_t)n + a];
                s_ap[2] = P[2 * (size_t)n + a];
                double h = R.chain_left, cb = INFINITY;
                int ac[3] = {0, 0, 0};
                if (cps) {
                    for (int q = 0; q < 3; q++) ac[q] = bcell1(s_ap[q], R.side, cps);
                    cb = A.bounds[R.bound_off + ac[d] + 1] - s_ap[d];
                    if (cb < h) h = cb;
                    int k = 0, acc = 0;
                    for (int dz = -1; dz <= 1; dz++)
                        for (int dy = -1; dy <= 1; dy++)
                            for (int dx = -1; dx <= 1; dx++) {
                                int cx = (ac[0] + dx + cps) % cps, cy = (ac[1] + dy + cps) % cps,
                                    cz = (ac[2] + dz + cps) % cps;
                                int c = cx + cps * (cy + cps * cz);
                                s_nearby[k] = c;
                                s_pref[k] = acc;
                                acc += min(cell_n[c], B_CELL_CAP);
                                k++;
                            }
                    s_pref[27] = acc;
                    s_M = acc;
                } else {
                    s_M = n - 1;
                }
                s_h = h;
                s_cb = cb;
                s_nc = 0;
                s_overflow = 0;
            }
            bsync();
            const int M = s_M;
            const double h = s_h;
            const double ap[3] = {s_ap[0], s_ap[1], s_ap[2]};
            const float apf[3] = {(float)ap[0], (float)ap[1], (float)ap[2]};
            const float hf = (float)h * (1.f + 0x1p-22f) + 0x1p-60f;
            BKey best = {INFINITY, 3, 0x7fffffff, 0};
            unsigned long long geo_pass = 0;
            if (A.mode != 1) {
                // ---- pass 1: FP32 lower bounds
                for (int idx = tid; idx < M; idx += NTH) {
                    int j;
                    if (cps) {
                        int c = 0;
                        while (s_pref[c + 1] <= idx) c++;
                        j = cell_l[(size_t)s_nearby[c] * B_CELL_CAP + (idx - s_pref[c])];
                    } else {
                        j = idx < a ? idx : idx + 1;
                    }
                    float sv[3] = {px[j] - apf[0], py[j] - apf[1], pz[j] - apf[2]};
                    bool ambiguous = false;
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        if (sv[q] >= 0.5f * Lf) sv[q] -= Lf;
                        else if (sv[q] < -0.5f * Lf) sv[q] += Lf;
                        if (fabsf(fabsf(sv[q]) - 0.5f * Lf) <= 2.f * E) ambiguous = true;
                    }
                    Geo g;
                    g.sd_lo = sv[d] - E;
                    g.sd_hi = sv[d] + E;
                    float p2lo = 0.f, p2hi = 0.f;
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        if (q == d) continue;
                        float m = fabsf(sv[q]);
                        float lo = fmaxf(m - E, 0.f), hi = m + E;
                        p2lo += lo * lo;
                        p2hi += hi * hi;
                    }
                    g.perp2_lo = dn(p2lo);
                    g.perp2_hi = up(p2hi);
                    float mdl = fmaxf(fabsf(sv[d]) - E, 0.f), mdh = fabsf(sv[d]) + E;
                    g.r2_lo = dn(p2lo + mdl * mdl);
                    g.r2_hi = up(p2hi + mdh * mdh);
                    float tl[2] = {0.f, 0.f};
                    bool any = false;
                    for (int f = 0; f < A.fac.n_factors; f++) {
                        float k = (float)A.fac.prefactor[f];
                        float fl = ambiguous ? 0.f
                                             : (g.perp2_lo >= rc2_hi ? __int_as_float(0x7f800000)
                                                                      : factor_geometric_floor(g, k, rc2_hi));
                        tl[f] = fl;
                        if (fl <= hf) any = true;
                    }
                    if (!any) continue;
                    geo_pass++;
                    uint32_t w[4];
                    philox4x32_10((uint32_t)leg, (uint32_t)(leg >> 32), (uint32_t)j, 0u, (uint32_t)r, A.seed_lo, w);
                    for (int f = 0; f < A.fac.n_factors; f++) {
                        if (tl[f] > hf) continue;
                        float t = 0.f;
                        if (!ambiguous) {
                            double u = unit53(w[2 * f], w[2 * f + 1]);
                            // -log(1-u)/beta with a 2^-40 downward margin over the FP64 log
                            double dU = -log(1.0 - u) / A.beta;
                            float dU_lo = (float)(dU * (1.0 - 0x1p-40)) * (1.f - 0x1p-22f);
                            t = factor_lower_bound(g, (float)A.fac.prefactor[f], (int)A.fac.power[f], rc2_lo, rc2_hi,
                                                   vrc_lo[f], vrc_hi[f], dU_lo);
                        }
                        if (t <= hf) {
                            int slot = atomicAdd(&s_nc, 1);
                            if (slot < B_CONTENDERS) cont_id[slot] = (j << 1) | f;
                            else s_overflow = 1;
                        }
                    }
                }
                bsync();
            }
            const bool all_exact = (A.mode == 1) || s_overflow;
            // ---- pass 2: exact FP64 inversion of the contenders (or of every candidate)
            const int nwork = all_exact ? M * A.fac.n_factors : s_nc;
            for (int w = tid; w < nwork; w += NTH) {
                int j, f;
                if (all_exact) {
                    int idx = w / A.fac.n_factors;
                    f = w % A.fac.n_factors;
                    if (cps) {
                        int c = 0;
                        while (s_pref[c + 1] <= idx) c++;
                        j = cell_l[(size_t)s_nearby[c] * B_CELL_CAP + (idx - s_pref[c])];
                    } else {
                        j = idx < a ? idx : idx + 1;
                    }
                } else {
                    j = cont_id[w] >> 1;
                    f = cont_id[w] & 1;
                }
                double dx = exact_displacement(A, R, ap, d, j, f, leg, (uint32_t)r);
                BKey c = {dx, 0, j, f};
                if (bkey_lt(c, best)) best = c;
            }
            if (tid == 0) {
                BKey c1 = {s_cb, 1, 0, 0};
                if (bkey_lt(c1, best)) best = c1;
                BKey c2 = {R.chain_left, 2, 0, 0};
                if (bkey_lt(c2, best)) best = c2;
            }
            block_min<WPR>(best, sk);
            if (A.mode == 2) {
                // audit: every candidate's FP32 bound must not exceed its exact displacement
                // when the latter is <= h (only those can decide the event)
                for (int idx = tid; idx < M * A.fac.n_factors; idx += NTH) {
                    int ci = idx / A.fac.n_factors, f = idx % A.fac.n_factors, j;
                    if (cps) {
                        int c = 0;
                        while (s_pref[c + 1] <= ci) c++;
                        j = cell_l[(size_t)s_nearby[c] * B_CELL_CAP + (ci - s_pref[c])];
                    } else {
                        j = ci < a ? ci : ci + 1;
                    }
                    double dx = exact_displacement(A, R, ap, d, j, f, leg, (uint32_t)r);
                    if (dx <= h) {
                        bool found = false;
                        for (int q = 0; q < min(s_nc, B_CONTENDERS); q++)
                            if (cont_id[q] == ((j << 1) | f)) found = true;
                        if (!found && !s_overflow) atomicAdd(A.audit, 1ull);
                    }
                }
            }
            if (tid == 0) {
                s_stats[0] += (unsigned long long)M;
                s_stats[2] += (unsigned long long)min(s_nc, B_CONTENDERS);
                s_stats[3] += (unsigned long long)(all_exact ? M * A.fac.n_factors : min(s_nc, B_CONTENDERS));
            }
            if (geo_pass) atomicAdd(&s_stats[1], geo_pass);
            // ---- commit
            if (tid == 0) {
                double *Pw = A.pos64 + R.off;
                double x;
                ecmc_batch_event ev;
                ev.kind = best.rank;
                ev.active = a;
                ev.target = -1;
                ev.factor = 0;
                ev.displacement = best.d;
                if (best.rank == 1) {
                    int c = bcell1(ap[d], R.side, cps);
                    x = c + 1 < cps ? A.bounds[R.bound_off + c + 1] : 0.0;
                } else {
                    x = ap[d] + best.d;
                    if (x >= L) x -= L;
                }
                Pw[(size_t)d * n + a] = x;
                float xf = (float)x;
                if (d == 0) px[a] = xf;
                else if (d == 1) py[a] = xf;
                else pz[a] = xf;
                R.est[1] += best.d;
                int newa = a;
                if (best.rank == 0) {
                    const int j = best.j;
                    double sdj = bmin_image(Pw[(size_t)d * n + j] - ap[d], L) - best.d;
                    R.est[0] += sdj;
                    R.est[2] += 1.0;
                    R.est[4] += 1.0;
                    R.chain_left -= best.d;
                    newa = j;
                    ev.target = j;
                    ev.factor = best.f;
                } else if (best.rank == 1) {
                    R.est[3] += 1.0;
                    R.chain_left -= best.d;
                } else {
                    uint32_t w[4];
                    philox4x32_10((uint32_t)leg, (uint32_t)(leg >> 32), 0xFFFFFFFFu, 1u, (uint32_t)r, A.seed_lo, w);
                    newa = (int)(((uint64_t)w[0] * (uint64_t)n) >> 32);
                    R.est[5] += 1.0;
                    R.chain_left = A.chain_length;
                    R.direction = (d + 1) % 3;
                    ev.target = newa;
                }
                ev.x = x;
                if (newa != a && cps) {
                    // the old active unit joins the cell it ended in, the new one leaves its cell
                    double pa[3] = {ap[0], ap[1], ap[2]};
                    pa[d] = x;
                    int ca = bcell1(pa[0], R.side, cps) + cps * (bcell1(pa[1], R.side, cps) + cps * bcell1(pa[2], R.side, cps));
                    int na = cell_n[ca];
                    if (na >= B_CELL_CAP) R.err = 1;
                    else { cell_l[(size_t)ca * B_CELL_CAP + na] = (int16_t)a; cell_n[ca] = na + 1; }
                    double pn[3] = {Pw[newa], Pw[(size_t)n + newa], Pw[2 * (size_t)n + newa]};
                    int cn = bcell1(pn[0], R.side, cps) + cps * (bcell1(pn[1], R.side, cps) + cps * bcell1(pn[2], R.side, cps));
                    int m = cell_n[cn], q = 0;
                    while (q < m && cell_l[(size_t)cn * B_CELL_CAP + q] != newa) q++;
                    if (q == m) R.err = 2;
                    else { cell_l[(size_t)cn * B_CELL_CAP + q] = cell_l[(size_t)cn * B_CELL_CAP + m - 1]; cell_n[cn] = m - 1; }
                }
                R.active = newa;
                if (A.log && leg < (uint64_t)A.log_legs) A.log[(size_t)r * A.log_legs + leg] = ev;
                R.leg = leg + 1;
            }
            bsync();
        }
        if (tid == 0) {
            for (int q = 0; q < 4; q++) R.stats[q] += (int64_t)s_stats[q];
            A.rep[r] = R;
        }
        bsync();
    }
}

__global__ void zero_counters(int32_t *q, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = 0;
}

}  // namespace ecmc

using namespace ecmc;

struct ecmc_batch {
    ecmc_batch_desc desc;
    int n_rep = 0;
    std::vector<int32_t> n;
    std::vector<ReplicaDev> rep_host;
    ReplicaDev *d_rep = nullptr;
    double *d_pos = nullptr, *d_bounds = nullptr;
    int64_t total_units = 0;
    // size classes
    int n_classes = 0;
    std::vector<int32_t> class_max_n, class_wpr, class_count, class_nmax, class_cellmax;
    std::vector<int32_t *> d_order;
    int32_t *d_queue = nullptr;
    std::vector<hipStream_t> streams;
    hipStream_t main_stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> fork;
    ecmc_batch_event *d_log = nullptr;
    int32_t log_legs = 0;
    unsigned long long *d_audit = nullptr;
    int32_t mode = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int32_t graph_legs = -1;
};

static size_t lds_bytes(int nmax, int cellmax) {
    size_t b = 12ull * nmax + 4ull * cellmax + 2ull * cellmax * B_CELL_CAP;
    b += b & 3;
    b += 4ull * B_CONTENDERS;
    return (b + 15) & ~15ull;
}

static int launch_class(ecmc_batch *b, int c, int legs, hipStream_t st) {
    if (b->class_count[c] == 0) return ECMC_OK;
    BatchArgs A;
    memset(&A, 0, sizeof(A));
    A.rep = b->d_rep;
    A.pos64 = b->d_pos;
    A.bounds = b->d_bounds;
    A.order = b->d_order[c];
    A.count = b->class_count[c];
    A.nmax = b->class_nmax[c];
    A.cellmax = b->class_cellmax[c];
    A.legs = legs;
    A.mode = b->mode;
    A.log_legs = b->log_legs;
    A.seed_lo = (uint32_t)b->desc.seed;
    A.beta = b->desc.beta;
    A.chain_length = b->desc.chain_length;
    A.fac.n_factors = b->desc.n_factors;
    for (int f = 0; f < 2; f++) { A.fac.power[f] = b->desc.power[f]; A.fac.prefactor[f] = b->desc.prefactor[f]; }
    A.queue = b->d_queue + c;
    A.log = b->d_log;
    A.audit = b->d_audit;
    size_t lds = lds_bytes(A.nmax, A.cellmax);
    int wpr = b->class_wpr[c];
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int per_cu = 1;
#define ECMC_LAUNCH(W)                                                                                     \
    {                                                                                                      \
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_kernel<W>, 64 * W, lds);              \
        if (per_cu < 1) return ecmc_fail(ECMC_ERR_CONFIGURATION, "replica does not fit in LDS");           \
        int grid = std::min(A.count, per_cu * std::max(cus, 1));                                           \
        hipLaunchKernelGGL(batched_kernel<W>, dim3(grid), dim3(64 * W), lds, st, A);                       \
    }
    switch (wpr) {
    case 1: ECMC_LAUNCH(1); break;
    case 2: ECMC_LAUNCH(2); break;
    case 4: ECMC_LAUNCH(4); break;
    case 8: ECMC_LAUNCH(8); break;
    default: return ecmc_fail(ECMC_ERR_ARGUMENT, "wavefronts per replica must be 1, 2, 4 or 8");
    }
#undef ECMC_LAUNCH
    if (hipGetLastError() != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "batched kernel launch failed");
    return ECMC_OK;
}

// One step: reset the work queues, then every size class on its own stream.
static int enqueue_step(ecmc_batch *b, int legs, hipStream_t main) {
    hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, main, b->d_queue, b->n_classes);
    hipEventRecord(b->fork[0], main);
    for (int c = 0; c < b->n_classes; c++) {
        hipStreamWaitEvent(b->streams[c], b->fork[0], 0);
        int rc = launch_class(b, c, legs, b->streams[c]);
        if (rc) return rc;
        hipEventRecord(b->fork[c + 1], b->streams[c]);
        hipStreamWaitEvent(main, b->fork[c + 1], 0);
    }
    return ECMC_OK;
}

extern "C" int ecmc_batch_create(const ecmc_batch_desc *desc, const int32_t *n_per_replica,
                                 const double *system_length, const double *positions, ecmc_batch **out) {
    if (!desc || !n_per_replica || !system_length || !positions || !out) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (desc->n_replicas < 1) return ecmc_fail(ECMC_ERR_ARGUMENT, "at least one replica");
    if (desc->n_factors < 1 || desc->n_factors > 2) return ecmc_fail(ECMC_ERR_CONFIGURATION, "one or two factors");
    for (int f = 0; f < desc->n_factors; f++) {
        double p = desc->power[f];
        if (!(p == 6.0 || p == 12.0)) return ecmc_fail(ECMC_ERR_CONFIGURATION, "factor powers 6 and 12 are supported");
        if (desc->prefactor[f] == 0.0) return ecmc_fail(ECMC_ERR_CONFIGURATION, "zero prefactor");
    }
    if (!(desc->cutoff > 0.0) || !(desc->beta > 0.0) || !(desc->chain_length > 0.0))
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "cutoff, beta and chain_length must be positive");
    ecmc_batch *b = new ecmc_batch();
    b->desc = *desc;
    b->n_rep = desc->n_replicas;
    b->n.assign(n_per_replica, n_per_replica + b->n_rep);
    std::vector<double> bounds;
    int64_t off = 0;
    b->rep_host.resize(b->n_rep);
    std::vector<double> soa;
    const double *pin = positions;
    for (int r = 0; r < b->n_rep; r++) {
        int nr = b->n[r];
        double L = system_length[r];
        if (nr < 2 || nr > 16384 || !(L > 0.0)) { delete b; return ecmc_fail(ECMC_ERR_CONFIGURATION, "replica size out of range"); }
        ReplicaDev &R = b->rep_host[r];
        memset(&R, 0, sizeof(R));
        R.n = nr;
        R.off = off;
        R.L = L;
        R.cutoff = std::min(desc->cutoff, L / 2.0);
        int cps = (int)(L / R.cutoff);
        if (cps >= 3) {
            if (cps > B_MAX_CPS) { delete b; return ecmc_fail(ECMC_ERR_CONFIGURATION, "more than 8 cells per side"); }
            R.cps = cps;
            R.side = L / cps;
            R.bound_off = (int32_t)bounds.size();
            for (int c = 0; c < cps; c++) {
                double lo = c * R.side;
                if (lo > 0.0) {
                    while ((int)(lo / R.side) == c) lo = nextafter(lo, -INFINITY);
                    while ((int)(lo / R.side) < c) lo = nextafter(lo, INFINITY);
                }
                bounds.push_back(lo);
            }
            bounds.push_back(L);
        } else {
            R.cps = 0;
            R.side = L;
            R.bound_off = 0;
        }
        R.active = 0;
        R.direction = 0;
        R.chain_left = desc->chain_length;
        R.leg = 0;
        for (int i = 0; i < nr; i++)
            for (int q = 0; q < 3; q++) {
                double x = pin[3 * i + q];
                if (!(x >= 0.0 && x < L)) { delete b; return ecmc_fail(ECMC_ERR_STATE, "initial position outside [0, L)"); }
            }
        soa.resize((size_t)(off + 3 * nr));
        for (int i = 0; i < nr; i++)
            for (int q = 0; q < 3; q++) soa[(size_t)off + (size_t)q * nr + i] = pin[3 * i + q];
        pin += 3 * (size_t)nr;
        off += 3 * nr;
    }
    b->total_units = off / 3;
    if (bounds.empty()) bounds.push_back(0.0);
    // default size classes (replaced by ecmc_batch_set_classes)
    b->n_classes = 0;
    int32_t def_max[3] = {256, 768, 16384}, def_wpr[3] = {1, 2, 4};
    int rc = 0;
    if (hipMalloc((void **)&b->d_rep, sizeof(ReplicaDev) * b->n_rep) != hipSuccess ||
        hipMalloc((void **)&b->d_pos, sizeof(double) * soa.size()) != hipSuccess ||
        hipMalloc((void **)&b->d_bounds, sizeof(double) * bounds.size()) != hipSuccess ||
        hipMalloc((void **)&b->d_queue, sizeof(int32_t) * 16) != hipSuccess ||
        hipMalloc((void **)&b->d_audit, sizeof(unsigned long long)) != hipSuccess) {
        ecmc_batch_destroy(b);
        return ecmc_fail(ECMC_ERR_HIP, "hipMalloc failed");
    }
    hipMemcpy(b->d_rep, b->rep_host.data(), sizeof(ReplicaDev) * b->n_rep, hipMemcpyHostToDevice);
    hipMemcpy(b->d_pos, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice);
    hipMemcpy(b->d_bounds, bounds.data(), sizeof(double) * bounds.size(), hipMemcpyHostToDevice);
    unsigned long long zero = 0;
    hipMemcpy(b->d_audit, &zero, sizeof(zero), hipMemcpyHostToDevice);
    hipStreamCreate(&b->main_stream);
    hipEventCreate(&b->ev0);
    hipEventCreate(&b->ev1);
    rc = ecmc_batch_set_classes(b, 3, def_max, def_wpr);
    if (rc) { ecmc_batch_destroy(b); return rc; }
    *out = b;
    return ECMC_OK;
}

extern "C" int ecmc_batch_set_classes(ecmc_batch *b, int32_t n_classes, const int32_t *class_max_n,
                                      const int32_t *class_wpr) {
    if (!b || n_classes < 1 || n_classes > 8 || !class_max_n || !class_wpr) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad classes");
    for (int c = 0; c < n_classes; c++) {
        int w = class_wpr[c];
        if (w != 1 && w != 2 && w != 4 && w != 8) return ecmc_fail(ECMC_ERR_ARGUMENT, "wavefronts per replica must be 1, 2, 4 or 8");
        if (c > 0 && class_max_n[c] <= class_max_n[c - 1]) return ecmc_fail(ECMC_ERR_ARGUMENT, "class bounds must increase");
    }
    if (class_max_n[n_classes - 1] < *std::max_element(b->n.begin(), b->n.end()))
        return ecmc_fail(ECMC_ERR_ARGUMENT, "the last class must hold the largest replica");
    for (auto p : b->d_order) hipFree(p);
    for (auto s : b->streams) hipStreamDestroy(s);
    for (auto e : b->fork) hipEventDestroy(e);
    b->d_order.clear(); b->streams.clear(); b->fork.clear();
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    b->n_classes = n_classes;
    b->class_max_n.assign(class_max_n, class_max_n + n_classes);
    b->class_wpr.assign(class_wpr, class_wpr + n_classes);
    b->class_count.assign(n_classes, 0);
    b->class_nmax.assign(n_classes, 2);
    b->class_cellmax.assign(n_classes, 0);
    std::vector<std::vector<int32_t>> members(n_classes);
    for (int r = 0; r < b->n_rep; r++) {
        int c = 0;
        while (b->n[r] > class_max_n[c]) c++;
        members[c].push_back(r);
    }
    for (int c = 0; c < n_classes; c++) {
        auto &m = members[c];
        std::stable_sort(m.begin(), m.end(), [&](int x, int y) { return b->n[x] > b->n[y]; });  // LPT
        b->class_count[c] = (int32_t)m.size();
        for (int r : m) {
            b->class_nmax[c] = std::max(b->class_nmax[c], b->n[r]);
            int cps = b->rep_host[r].cps;
            b->class_cellmax[c] = std::max(b->class_cellmax[c], cps * cps * cps);
        }
        int32_t *d = nullptr;
        if (hipMalloc((void **)&d, sizeof(int32_t) * std::max<size_t>(m.size(), 1)) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "hipMalloc failed");
        if (!m.empty()) hipMemcpy(d, m.data(), sizeof(int32_t) * m.size(), hipMemcpyHostToDevice);
        b->d_order.push_back(d);
        hipStream_t s;
        hipStreamCreate(&s);
        b->streams.push_back(s);
    }
    for (int c = 0; c <= n_classes; c++) {
        hipEvent_t e;
        hipEventCreateWithFlags(&e, hipEventDisableTiming);
        b->fork.push_back(e);
    }
    return ECMC_OK;
}

extern "C" void ecmc_batch_destroy(ecmc_batch *b) {
    if (!b) return;
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (auto p : b->d_order) hipFree(p);
    for (auto s : b->streams) hipStreamDestroy(s);
    for (auto e : b->fork) hipEventDestroy(e);
    hipFree(b->d_rep); hipFree(b->d_pos); hipFree(b->d_bounds); hipFree(b->d_queue); hipFree(b->d_log);
    hipFree(b->d_audit);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->main_stream) hipStreamDestroy(b->main_stream);
    delete b;
}

extern "C" int ecmc_batch_set_mode(ecmc_batch *b, int32_t mode) {
    if (!b || mode < 0 || mode > 2) return ecmc_fail(ECMC_ERR_ARGUMENT, "mode must be 0, 1 or 2");
    b->mode = mode;
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    return ECMC_OK;
}

extern "C" int ecmc_batch_set_log(ecmc_batch *b, int32_t legs_per_replica) {
    if (!b || legs_per_replica < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    hipFree(b->d_log);
    b->d_log = nullptr;
    b->log_legs = legs_per_replica;
    if (legs_per_replica > 0 &&
        hipMalloc((void **)&b->d_log, sizeof(ecmc_batch_event) * (size_t)legs_per_replica * b->n_rep) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMalloc of the log failed");
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    return ECMC_OK;
}

extern "C" int ecmc_batch_run(ecmc_batch *b, int32_t legs_per_replica) {
    if (!b || legs_per_replica < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    hipEventRecord(b->ev0, b->main_stream);
    int rc = enqueue_step(b, legs_per_replica, b->main_stream);
    if (rc) return rc;
    hipEventRecord(b->ev1, b->main_stream);
    if (hipStreamSynchronize(b->main_stream) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "batched kernel failed");
    return ECMC_OK;
}

// Capture one step (queue reset + every class) into a hipGraph, then replay it n_steps times.
extern "C" int ecmc_batch_run_graph(ecmc_batch *b, int32_t legs_per_replica, int32_t n_steps) {
    if (!b || legs_per_replica < 0 || n_steps < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (!b->graph_exec || b->graph_legs != legs_per_replica) {
        if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
        if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
        if (hipStreamBeginCapture(b->main_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "stream capture failed");
        int rc = enqueue_step(b, legs_per_replica, b->main_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->main_stream, &g);
        if (rc) { if (g) hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "stream capture failed");
        b->graph = g;
        if (hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "graph instantiation failed");
        b->graph_legs = legs_per_replica;
    }
    hipEventRecord(b->ev0, b->main_stream);
    for (int s = 0; s < n_steps; s++)
        if (hipGraphLaunch(b->graph_exec, b->main_stream) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "graph launch failed");
    hipEventRecord(b->ev1, b->main_stream);
    if (hipStreamSynchronize(b->main_stream) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "batched graph failed");
    return ECMC_OK;
}

extern "C" int ecmc_batch_last_ms(const ecmc_batch *b, double *ms) {
    if (!b || !ms) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    float f = 0.f;
    if (hipEventElapsedTime(&f, b->ev0, b->ev1) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "event timing failed");
    *ms = f;
    return ECMC_OK;
}

static int pull_replicas(const ecmc_batch *b, std::vector<ReplicaDev> &rep) {
    rep.resize(b->n_rep);
    if (hipMemcpy(rep.data(), b->d_rep, sizeof(ReplicaDev) * b->n_rep, hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    for (int r = 0; r < b->n_rep; r++)
        if (rep[r].err) return ecmc_fail(ECMC_ERR_STATE, "a cell list overflowed (more than 64 units in one cell)");
    return ECMC_OK;
}

extern "C" int ecmc_batch_states(const ecmc_batch *b, int32_t *active, int32_t *direction, double *chain_left,
                                 uint64_t *leg, double *est8, int64_t *stats4) {
    if (!b) return ecmc_fail(ECMC_ERR_ARGUMENT, "null batch");
    std::vector<ReplicaDev> rep;
    int rc = pull_replicas(b, rep);
    if (rc) return rc;
    for (int r = 0; r < b->n_rep; r++) {
        if (active) active[r] = rep[r].active;
        if (direction) direction[r] = rep[r].direction;
        if (chain_left) chain_left[r] = rep[r].chain_left;
        if (leg) leg[r] = rep[r].leg;
        if (est8) memcpy(est8 + 8 * (size_t)r, rep[r].est, sizeof(double) * 8);
        if (stats4) memcpy(stats4 + 4 * (size_t)r, rep[r].stats, sizeof(int64_t) * 4);
    }
    return ECMC_OK;
}

extern "C" int ecmc_batch_positions(const ecmc_batch *b, double *positions) {
    if (!b || !positions) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    std::vector<double> soa((size_t)3 * b->total_units);
    if (hipMemcpy(soa.data(), b->d_pos, sizeof(double) * soa.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    double *o = positions;
    for (int r = 0; r < b->n_rep; r++) {
        const ReplicaDev &R = b->rep_host[r];
        for (int i = 0; i < R.n; i++)
            for (int q = 0; q < 3; q++) *o++ = soa[(size_t)R.off + (size_t)q * R.n + i];
    }
    return ECMC_OK;
}

extern "C" int ecmc_batch_log(const ecmc_batch *b, ecmc_batch_event *host_log) {
    if (!b || !host_log || !b->d_log) return ecmc_fail(ECMC_ERR_ARGUMENT, "no log configured");
    if (hipMemcpy(host_log, b->d_log, sizeof(ecmc_batch_event) * (size_t)b->log_legs * b->n_rep,
                  hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    return ECMC_OK;
}

extern "C" int ecmc_batch_audit(const ecmc_batch *b, uint64_t *violations) {
    if (!b || !violations) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    unsigned long long v = 0;
    if (hipMemcpy(&v, b->d_audit, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    *violations = v;
    return ECMC_OK;
}
