// ecmc_fp32_bounds.h -- FP32 brackets of inverse-power event displacements, for the
// candidate filters of the batched kernel and of the chain kernel.
//
// Every quantity is bracketed [lo, hi] and each FP32 operation is covered by a relative
// margin (ETA for products/quotients/roots, ETA_POW or a logarithm-scaled margin for the
// exp2/log2 powers), far above FP32 rounding (2^-24 per operation) and the accuracy of the
// device's log2f and exp2f (each measured over every float of its domain by the tests: log2f
// within 2^-21 max(|log2 x|, 1), exp2f within 2^-21 relative).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ecmc {

// Every quantity is bracketed [lo, hi] and each FP32 operation is covered by a relative
// margin (ETA for products/quotients/roots, ETA_POW for the exp2/log2 power), far above
// FP32 rounding (2^-24 per operation) and the documented accuracy of log2f/exp2f.
constexpr float ETA = 0x1p-17f, ETA_POW = 0x1p-14f;
// x moved outward by ETA |x| with one fused operation: the single rounding leaves a margin of
// at least ETA - 2^-23 relative.  No argument is -inf (every caller passes a finite value or
// +inf, which stays +inf).
__device__ __forceinline__ float up(float x) { return fmaf(fabsf(x), ETA, x); }
__device__ __forceinline__ float dn(float x) { return fmaf(-fabsf(x), ETA, x); }
__device__ __forceinline__ float sq_lo(float a_abs_lo) { return dn(a_abs_lo * a_abs_lo); }
// Hardware reciprocal and square root (v_rcp_f32, v_sqrt_f32: within 1 ulp = 2^-23 relative,
// far inside the ETA margins); below 2^-100 the root is bracketed by [0, 2^-50] directly.
__device__ __forceinline__ float rcp1(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_up(float x) { return x > 0x1p-100f ? up(__builtin_amdgcn_sqrtf(x)) : 0x1p-50f; }
__device__ __forceinline__ float sqrt_dn(float x) { return x > 0x1p-100f ? dn(__builtin_amdgcn_sqrtf(x)) : 0.f; }

// The device log2f is within LOG2F_ERR * max(|log2 x|, 1) of log2 x for every positive float:
// an exhaustive scan (ecmc_fp32_math_error) measures 2^-23 and the tests require <= 2^-21.
constexpr float LOG2F_ERR = 0x1p-20f;

// FP32 lower bound of dU = -log(1 - u) / beta for u = unit53(a, b), from the first word
// alone: u >= u24 = (a >> 8) 2^-24 exactly, -log(1 - u) >= u, and for u24 >= 2^-10 the
// float 1 - u24 is exact and -log2f of it is within LOG2F_ERR max(|.|, 1) of
// -log2(1 - u24).  inv_beta_lo is a lower bound of 1 / beta.  The result never exceeds the
// exact dU.
__device__ __forceinline__ float dU_lower(uint32_t a, float inv_beta_lo) {
    const float u24 = (float)(a >> 8) * 0x1p-24f;
    float l;
    if (u24 < 0x1p-10f) {
        l = u24;                                           // -ln(1 - u) >= u >= u24
    } else {
        const float ml2 = -log2f(1.f - u24);
        const float ml2_lo = ml2 - LOG2F_ERR * fmaxf(ml2, 1.f);  // <= -log2(1 - u24) <= -log2(1 - u)
        l = ml2_lo * 0x1.62e42ep-1f;                       // ln 2 rounded down
    }
    return fmaxf(l * inv_beta_lo * (1.f - 0x1p-20f), 0.f);
}


// FP32 upper bound of dU = -log(1 - u) / beta for u = unit53(a, b): u < uh = ((a >> 8) + 1)
// 2^-24 (exact in float, +inf when uh reaches 1); -log(1 - u) <= u / (1 - u) <= uh (1 + 2^-9)
// for uh <= 2^-10, else -log2f(1 - uh) (1 - uh exact) plus its error bound, times ln 2
// rounded up.  inv_beta_hi is an upper bound of 1 / beta.
__device__ __forceinline__ float dU_upper(uint32_t a, float inv_beta_hi) {
    const uint32_t k = (a >> 8) + 1u;
    if (k >= (1u << 24)) return __int_as_float(0x7f800000);
    const float uh = (float)k * 0x1p-24f;
    float l;
    if (uh <= 0x1p-10f) {
        l = up(uh * (1.f + 0x1p-9f));
    } else {
        const float ml2 = -log2f(1.f - uh);
        const float ml2_hi = ml2 + LOG2F_ERR * fmaxf(ml2, 1.f);
        l = ml2_hi * 0x1.62e430p-1f;                       // ln 2 rounded up
    }
    return up(l * inv_beta_hi);
}

// K / r2^(p/2) for an integer power 1 <= p <= 32, bounded up (hi) or down: r2^(p/2) by p/2 - 1
// products and, for odd p, one hardware square root (1 ulp), then K rcp(.) (1 ulp and one
// product): at most (p/2 + 6) 2^-24 <= 22 2^-24 relative with K's own conversion, inside the
// factor 1 +- (2^-17 + 2^-19).  ok turns false outside the normal range (the caller then keeps
// the candidate).
__device__ __forceinline__ float vpow_g(float K, float r2, int p, bool hi, bool &ok) {
    if (!(r2 >= 0x1p-100f && r2 <= 0x1p100f)) { ok = false; return 0.f; }
    float rp = (p & 1) ? __builtin_amdgcn_sqrtf(r2) : 1.f;
    for (int i = 0; i < (p >> 1); i++) rp *= r2;
    if (!(rp >= 0x1p-120f && rp <= 0x1p120f)) { ok = false; return 0.f; }
    const float v = K * rcp1(rp);
    if (!(v >= 0x1p-120f && v <= 0x1p120f)) { ok = false; return 0.f; }
    return v * (hi ? 1.f + 0x1p-17f + 0x1p-19f : 1.f - 0x1p-17f - 0x1p-19f);
}

// (K / V)^(2/p) bounded up (hi) or down, as exp2f(two_over_p log2f(q)), q = K rcp(V): the error
// of the exponent (log2f's 2^-21 max(|log2 q|, 1), q's 2^-22, two products) stays below
// 2^-19.5 (max(|log2 q|, 1) + 1), which with exp2f's 2^-21 is covered 4 times over by the
// margin 2^-17 (max(|log2 q|, 1) + 1).
__device__ __forceinline__ float rstar2_g(float K, float V, float two_over_p, bool hi, bool &ok) {
    if (!(V >= 0x1p-120f && V <= 0x1p120f)) { ok = false; return 0.f; }
    const float q = K * rcp1(V);
    if (!(q >= 0x1p-120f && q <= 0x1p120f)) { ok = false; return 0.f; }
    const float l2 = log2f(q);
    const float t = two_over_p * l2;
    if (!(t > -100.f && t < 100.f)) { ok = false; return 0.f; }
    const float r = exp2f(t);
    const float m = 0x1p-17f * (fmaxf(fabsf(l2), 1.f) + 1.f);
    return hi ? r * (1.f + m) : r * (1.f - m);
}

// Brackets [lb, ub] of the event time displacement of one two-leaf inverse-power factor
// (potential/inverse_power_potential.py:96-179 through ip_displacement, and its truncated-and-
// shifted form ipt_displacement): separation s (FP64, target minus active, minimum image),
// motion along +dir, K = |prefactor c1 c2| (> 0), rep = the factor repels, p its integer power,
// trunc with the squared cutoff bracketed by [rc2_lo, rc2_hi], dU bracketed by [dU_lo, dU_hi],
// 1 / speed by [isp_lo, isp_hi]; s may differ from the exact FP64 separation by abs_slack per
// component.  lb = +inf only where the exact displacement is +inf (no
// event); ub = +inf where an event is not certain.  Every case the brackets cannot decide gives
// lb = 0, ub = +inf.  A positive lb bounds the FP64 displacement of ip_displacement /
// ipt_displacement from below; where lb = 0 the FP64 value can be a rounding-level negative
// number (an event at the start, computed as s_d - sqrt(a) with a = s_d^2 in exact terms).
template <int D>
__device__ void ip_time_bounds(const double *s, int dir, float K, bool rep, int p, float two_over_p, bool trunc,
                               float rc2_lo, float rc2_hi, float dU_lo, float dU_hi, float isp_lo, float isp_hi,
                               float &lb, float &ub, float abs_slack = 0.f) {
    const float INF = __int_as_float(0x7f800000);
    // absolute slack per separation component: a component converted below the normal range, and
    // the caller's own (a separation within abs_slack of the exact FP64 one)
    const float TINY = fmaxf(abs_slack, 0x1p-110f);
    lb = 0.f;
    ub = INF;
    float lo2 = 0.f, hi2 = 0.f;     // the squared perpendicular components, bounded
#pragma unroll
    for (int i = 0; i < D; i++) {
        if (i == dir) continue;
        const float m = fabsf((float)s[i]);
        const float lo = fmaxf(dn(m) - TINY, 0.f), hi = up(m) + TINY;
        lo2 += lo * lo;
        hi2 += hi * hi;
    }
    const float perp2_lo = dn(lo2), perp2_hi = up(hi2);
    const float sdf = (float)s[dir];
    const float sd_lo = dn(sdf) - TINY, sd_hi = up(sdf) + TINY;
    if (trunc && perp2_lo >= rc2_hi) { lb = INF; return; }
    if (sd_lo <= 0.f && sd_hi > 0.f) return;          // the sign along the motion is not certain
    bool ok = true;
    // the potential's magnitude at the cutoff (truncated form), bracketed
    const float U_lo = trunc ? vpow_g(K, rc2_hi, p, false, ok) : 0.f;
    const float U_hi = trunc ? vpow_g(K, rc2_lo, p, true, ok) : 0.f;
    float l = 0.f, u = INF;
    if (rep) {
        if (sd_hi <= 0.f) { lb = INF; return; }     // behind: no event
        const float r2_lo = dn(lo2 + sd_lo * sd_lo), r2_hi = up(hi2 + sd_hi * sd_hi);
        // cur = U(min(r2, rc2)), maxp = U(perp2)
        const float cur_lo = vpow_g(K, trunc ? fminf(r2_hi, rc2_hi) : r2_hi, p, false, ok);
        const float cur_hi = vpow_g(K, trunc ? fminf(r2_lo, rc2_lo) : r2_lo, p, true, ok);
        const float mx_hi = perp2_lo >= 0x1p-100f ? vpow_g(K, perp2_lo, p, true, ok) : INF;
        if (!ok) return;
        if (dU_lo >= up(mx_hi - cur_lo)) { lb = INF; return; }   // the potential cannot rise by dU
        bool okl = true;
        const float ns_hi = rstar2_g(K, dn(cur_lo + dU_lo), two_over_p, true, okl);
        if (okl) l = fmaxf(dn(sd_lo - sqrt_up(fmaxf(up(ns_hi - perp2_lo), 0.f))), 0.f);
        if ((!trunc || perp2_hi < rc2_lo) && perp2_hi >= 0x1p-100f) {
            bool oku = true;
            const float mx_lo = vpow_g(K, perp2_hi, p, false, oku);
            if (oku && dU_hi < dn(mx_lo - cur_hi)) {
                const float ns_lo = rstar2_g(K, up(cur_hi + dU_hi), two_over_p, false, oku);
                if (oku) u = up(sd_hi - sqrt_dn(fmaxf(dn(ns_lo - perp2_hi), 0.f)));
            }
        }
    } else {
        // a unit ahead is first passed (displacement s_d to the closest approach); then the
        // separation s' (s_d' = min(s_d, 0)) grows and the potential -W rises towards the floor
        const bool ahead = sd_lo > 0.f;
        const float d0_lo = ahead ? sd_lo : 0.f, d0_hi = ahead ? sd_hi : 0.f;
        const float sp_lo = ahead ? 0.f : sd_lo, sp_hi = ahead ? 0.f : sd_hi;
        const float s2_lo = ahead ? perp2_lo : dn(lo2 + sd_hi * sd_hi);
        const float s2_hi = ahead ? perp2_hi : up(hi2 + sd_lo * sd_lo);
        if (trunc && s2_lo >= rc2_hi) { lb = INF; return; }
        const float W_hi = s2_lo >= 0x1p-100f ? vpow_g(K, s2_lo, p, true, ok) : INF;
        if (!ok) return;
        if (dU_lo >= up(W_hi - U_lo)) { lb = INF; return; }     // the potential cannot rise by dU
        bool okl = true;
        const float ns_lo = rstar2_g(K, up(W_hi - dU_lo), two_over_p, false, okl);
        const float x_lo = okl ? sqrt_dn(fmaxf(dn(ns_lo - perp2_hi), 0.f)) : 0.f;
        l = fmaxf(dn(dn(d0_lo + sp_lo) + x_lo), 0.f);
        if ((!trunc || s2_hi < rc2_lo) && s2_hi >= 0x1p-100f) {
            bool oku = true;
            const float W_lo = vpow_g(K, s2_hi, p, false, oku);
            const float den_lo = dn(W_lo - dU_hi);
            if (oku && den_lo > U_hi) {
                const float ns_hi = rstar2_g(K, den_lo, two_over_p, true, oku);
                if (oku) u = up(up(d0_hi + sp_hi) + sqrt_up(fmaxf(up(ns_hi - perp2_lo), 0.f)));
            }
        }
    }
    lb = fmaxf(dn(l * isp_lo), 0.f);
    ub = u == INF ? INF : up(u * isp_hi);
}

// FP32 brackets [lb, ub] of the time displacement of the inverse-power Coulomb bound
// (ipcb_displacement, inverse_power_coulomb_bounding_potential.c:85-139) with prefactor magnitude
// in [K_lo, K_hi], the squared perpendicular separation in [perp2_lo, perp2_hi], box length L
// (half = L / 2) and potential change in [dU_lo, dU_hi].  The displacement is floor(dU / per) L
// for the whole windings of the image sum, per = |pp| (1 / rho - 1 / sqrt(half^2 + rho^2)) the
// rise over one winding, plus a remainder in [0, 3L/2]: the way to the start of the rising
// flank (at most L) and the rise within it (at most half).  So lb = floor(q_lo) L and
// ub = (floor(q_hi) + 3/2) L for q_lo <= dU / per <= q_hi, each with relative margins far above
// the FP64 rounding of the exact inversion (2^-48 for per, whose cancellation is bounded since
// rho <= half sqrt(2)).  An event is always certain to come, so lb stays finite (capped at 2^100,
// which culls like any large bound).
__device__ __forceinline__ void ipcb_time_bounds(float K_lo, float K_hi, float perp2_lo, float perp2_hi, float L,
                                                 float dU_lo, float dU_hi, float isp_lo, float isp_hi, float &lb,
                                                 float &ub) {
    const float INF = __int_as_float(0x7f800000);
    const float h = 0.5f * L;
    const float h2_lo = dn(h * h), h2_hi = up(h * h);
    // a = 1 / rho and b = 1 / sqrt(half^2 + rho^2), bracketed
    const float a_hi = perp2_lo > 0x1p-100f ? up(rcp1(sqrt_dn(perp2_lo))) : INF;
    const float a_lo = dn(rcp1(sqrt_up(perp2_hi)));
    const float b_hi = up(rcp1(sqrt_dn(dn(h2_lo + perp2_lo))));
    const float b_lo = dn(rcp1(sqrt_up(up(h2_hi + perp2_hi))));
    const float d_lo = dn(a_lo - b_hi), d_hi = a_hi == INF ? INF : up(a_hi - b_lo);
    const float per_lo = d_lo > 0.f ? dn(dn(K_lo * d_lo) * (1.f - 0x1p-20f)) : 0.f;
    const float per_hi = d_hi == INF ? INF : up(up(K_hi * d_hi) * (1.f + 0x1p-20f));
    const float q_lo = per_hi == INF ? 0.f : dn(dU_lo * dn(rcp1(per_hi)));
    const float q_hi = per_lo > 0x1p-100f ? up(dU_hi * up(rcp1(per_lo))) : INF;
    const float k_lo = floorf(q_lo), k_hi = floorf(q_hi);
    const float l = k_lo > 0.f ? fminf(dn(dn(k_lo * L) * (1.f - 0x1p-20f)), 0x1p100f) : 0.f;
    const float u = k_hi < 0x1p100f ? up(up((k_hi + 1.5f) * L) * (1.f + 0x1p-20f)) : INF;
    lb = fminf(fmaxf(dn(l * isp_lo), 0.f), 0x1p100f);
    ub = u == INF ? INF : up(u * isp_hi);
}

}  // namespace ecmc
