// ecmc_chain.hip -- one ECMC chain advanced on the GPU, leg after leg, in one persistent
// workgroup (the plain form).
//
// A leg follows the reference mediator (mediator/single_process_mediator.py:84-156):
//   update the cell occupancy for the active unit, run the event handlers of the taggers in
//   the `create` list of the tagger that committed last (tag_activator.py:228-289), take the
//   earliest stored candidate, build and commit its out-state, and trash by tag
//   (tag_activator.py:291-316).
//
// Mapping to the hardware:
//   * the candidate set of a pair tagger (active unit x occupants of the nearby cells,
//     excluded_cells_tagger.py:129-132) is spread over the 256 threads of the workgroup in
//     canonical order (nearby cells in itertools.product order, occupants in list order),
//     one candidate per lane; positions (SoA x[], y[], z[]), the random ring, charges and cell
//     lists are read from LDS where the launch placed them, else from HBM;
//   * each candidate's exponential draw is the 2 MT19937 words at its rank in the stream;
//     words come from a ring that the workgroup refills 624 words at a time with the
//     three-phase parallel twist, so the draws are those of CPython's random in the
//     reference's call order;
//   * the inverse-power candidates are first bracketed in FP32 and only those that can be the
//     earliest are inverted exactly in FP64;
//   * the earliest candidate is found by a wavefront shuffle min-reduce over (time, push
//     rank) followed by a 4-wave LDS reduction; the scheduler keeps, per tagger, the
//     earliest of its pushed events (all events of one tagger are trashed together), and
//     ties go to the first pushed event as in list_scheduler.py:129-156; the earliest stored
//     event is chosen by a butterfly over the first wavefront's lanes, one slot per lane;
//   * bookkeeping that is sequential in the reference (occupancy lists, end of chain,
//     sampling, out-states) runs on thread 0 between barriers;
//   * the leg loop is compiled once per feature set (FEAT_*) and each launch runs the smallest
//     variant that covers its systems.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <string>
#include <mutex>
#include <utility>
#include <cmath>

#include "../../include/jellyfysh_amd.h"
#include "ecmc_math.h"
#include "ecmc_potentials.h"
#include "ecmc_fp32_bounds.h"
#include "ecmc_common.h"

namespace ecmc {

constexpr int NT = 256;

// The dynamic LDS of a leg-loop launch (positions, then the ring of tempered words), when the
// launch's arguments place them there (ChainArgs::lds_pos, lds_ring).
extern __shared__ __align__(16) unsigned char chain_dyn_lds[];

struct ChainArgs;
__device__ __forceinline__ int32_t *occ_base(const ChainArgs &a);
__device__ __forceinline__ int32_t *occn_base(const ChainArgs &a);
constexpr uint32_t MT_UPPER = 0x80000000u, MT_LOWER = 0x7fffffffu, MT_MATRIX = 0x9908b0dfu;

struct Slot {
    Time t;
    int64_t seq;
    int32_t valid, target;
    double apos[3];
    Time astamp;
    double tpos[3];
    double cb_boundary;
    int32_t cb_dir, adir;
    double aspeed;
    int32_t cb_cell, pad_cb;   // active cell of a cell-bounding candidate; target cell of a cell veto
    double cv_rate;            // cell veto: bounding event rate of the sampled cell
    int32_t cv_index, cv_phase;  // cell veto: drawn table entry; 1 = entry pending, 2 = entry drawn
};

// One entry of Walker's table (walker.py:73-101): item0 with rate0, else the alias item1
struct CvEntry {
    double rate0;
    int32_t item0, item1;
};

struct ChainState {
    int32_t active, adir;
    double aspeed;
    Time at;
    int32_t occ_id, occ_cell;
    int32_t started, done, err, pending_creates, last_tagger, stop_reason;
    Time last;
    int64_t seq;
    Time eoc_last, samp_time;
    uint64_t cursor, produced;
    double est[8];
    int32_t ncand;
    uint32_t chash;
    int32_t nsur, pad_s;       // surplus units of a bounded occupancy
    Slot slot[ECMC_MAX_TAGGERS];
};

struct ChainArgs {
    int32_t n, ncells, nb_stride, occ_cap, ring_words, n_taggers, stop_mask, ewald_fc;
    int32_t leaves;            // leaf units per composite object (1: point masses)
    int32_t occupancy;         // some tagger reads the cell occupancy (else it is not kept up)
    double L, half, beta;
    int32_t cps[3], cum[3];
    double side[3];
    Ewald ewald;
    double *pos;
    const double *charge;
    const double *cmin;
    const int32_t *nearby;
    const int32_t *nearby_n;
    int32_t *occ;
    int32_t *occ_n;
    uint32_t *mt;
    uint32_t *ring;
    ChainState *st;
    const ecmc_tagger_desc *tg;
    ecmc_event *log;
    int64_t max_events;
    int64_t *n_out;
    // bounded occupancy (single_active_cell_occupancy.py:62) and cell bounding
    int32_t max_occ, sur_cap, layers, pad2;
    int32_t *sur_cell, *sur_id;   // surplus units in the order the reference's dict yields them
    int32_t *cand;                // the cell-bounding tagger's candidates: their far cells (one occupant each)
    const double *cbt;            // cell-bound table: (cell offset, direction) -> upper, lower
    // cell veto: Walker tables [direction * 2 + (0 upper, 1 lower)] of cv_m entries each
    const CvEntry *cvw;
    double cv_total[6], cv_mean[6];
    int32_t cv_m, pad3;
    // residency in the launch's dynamic LDS: positions (dim x n doubles at offset 0) and the MT
    // ring (ring_words words at lds_ring_off), copied in at the start of a launch and back at
    // its end; 0 keeps them in global memory
    int32_t lds_pos, lds_ring, lds_ring_off, lds_cand_off;   // lds_cand_off < 0: the candidate list in global memory
    int32_t lds_occ_off, lds_occn_off;                         // < 0: the occupant lists / counts in global memory
    // the terms of the merged-image Coulomb derivative in the sequential loop's order (real-space
    // images by k, j, i, then Fourier terms by i, j, k): kind << 24 | (i + 128) << 16 |
    // (j + 128) << 8 | (k + 128), built on the host
    const int32_t *ewald_codes;
    int32_t ewald_ncodes, lds_charge_off;   // lds_charge_off < 0: the charges in global memory
};

// The cell occupant lists (ncells x occ_cap) and their counts, in LDS for a launch when placed there.
__device__ __forceinline__ int32_t *occ_base(const ChainArgs &a) {
    return a.lds_occ_off >= 0 ? reinterpret_cast<int32_t *>(chain_dyn_lds + a.lds_occ_off) : a.occ;
}
// The charge of unit id (constant over a run): from LDS when the launch placed the charges there.
__device__ __forceinline__ double charge_of(const ChainArgs &a, int id) {
    if (a.lds_charge_off >= 0) return reinterpret_cast<const double *>(chain_dyn_lds + a.lds_charge_off)[id];
    return a.charge[id];
}
__device__ __forceinline__ int32_t *occn_base(const ChainArgs &a) {
    return a.lds_occn_off >= 0 ? reinterpret_cast<int32_t *>(chain_dyn_lds + a.lds_occn_off) : a.occ_n;
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t y) {
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

// One twist of the 624-word state in LDS, then the tempered words go to the ring.
// The three phases follow the data dependences of the sequential loop: words 0..226 read
// only old words; 227..453 read new words 0..226; 454..623 read new words 227..396.
// Word j of the block goes to ring[(produced + j) & mask] (mask: ring size - 1, all ones for no wrap).
__device__ __forceinline__ void mt_generate_block(uint32_t *mt, uint32_t *ring, uint64_t mask, uint64_t produced) {
    const int t = threadIdx.x;
    uint32_t v = 0;
    if (t < 227) {
        uint32_t y = (mt[t] & MT_UPPER) | (mt[t + 1] & MT_LOWER);
        v = mt[t + 397] ^ (y >> 1) ^ ((y & 1u) ? MT_MATRIX : 0u);
    }
    __syncthreads();
    if (t < 227) mt[t] = v;
    __syncthreads();
    if (t < 227) {
        int i = t + 227;
        uint32_t y = (mt[i] & MT_UPPER) | (mt[i + 1] & MT_LOWER);
        v = mt[i - 227] ^ (y >> 1) ^ ((y & 1u) ? MT_MATRIX : 0u);
    }
    __syncthreads();
    if (t < 227) mt[t + 227] = v;
    __syncthreads();
    if (t < 170) {
        int i = t + 454;
        uint32_t nx = i == 623 ? mt[0] : mt[i + 1];
        uint32_t y = (mt[i] & MT_UPPER) | (nx & MT_LOWER);
        v = mt[i - 227] ^ (y >> 1) ^ ((y & 1u) ? MT_MATRIX : 0u);
    }
    __syncthreads();
    if (t < 170) mt[t + 454] = v;
    __syncthreads();
    for (int j = t; j < 624; j += NT) ring[(produced + (uint64_t)j) & mask] = mt_temper(mt[j]);
    __syncthreads();
}

// Collective: make `need` unread words available in the ring.
__device__ __forceinline__ void ring_ensure(ChainState &st, uint32_t *mt, const ChainArgs &a, uint64_t need) {
    if (need + 2 * 624 > (uint64_t)a.ring_words) {   // the ring would overwrite unread words
        st.err = ECMC_ERR_STATE;
        return;
    }
    while (st.produced < st.cursor + need) {
        mt_generate_block(mt, a.lds_ring ? reinterpret_cast<uint32_t *>(chain_dyn_lds + a.lds_ring_off) : a.ring,
                          (uint64_t)(a.ring_words - 1), st.produced);
        if (threadIdx.x == 0) st.produced += 624;
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t ring_word(const ChainArgs &a, uint64_t idx) {
    const uint64_t i = idx & (uint64_t)(a.ring_words - 1);   // ring_words is a power of two
    if (a.lds_ring) return reinterpret_cast<const uint32_t *>(chain_dyn_lds + a.lds_ring_off)[i];
    return a.ring[i];
}

// random() from two words (genrand_res53)
__device__ __forceinline__ double words_to_random(uint32_t w0, uint32_t w1) {
    uint32_t x = w0 >> 5, y = w1 >> 6;
    return (x * 67108864.0 + y) * (1.0 / 9007199254740992.0);
}

// candidate-set hash term: order dependent, summed over the candidates of a leg
__device__ __forceinline__ uint32_t cand_mix(uint32_t j, uint32_t target) {
    uint32_t h = (j * 0x9E3779B9u) ^ target;
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}

template <int D>
__device__ __forceinline__ int pos_to_cell(const ChainArgs &a, const double *p) {
    long idx = 0;
#pragma unroll
    for (int d = 0; d < D; d++) idx += (long)((int)(p[d] / a.side[d])) * a.cum[d];
    if (idx < 0 || idx >= a.ncells) return -1;
    return (int)idx;
}

template <int D>
__device__ __forceinline__ void load_pos(const ChainArgs &a, int id, double *p) {
    if (a.lds_pos) {
        const double *lp = reinterpret_cast<const double *>(chain_dyn_lds);
#pragma unroll
        for (int d = 0; d < D; d++) p[d] = lp[d * a.n + id];
    } else {
#pragma unroll
        for (int d = 0; d < D; d++) p[d] = a.pos[(size_t)d * a.n + id];
    }
}
template <int D>
__device__ __forceinline__ void store_pos(const ChainArgs &a, int id, const double *p) {
    if (a.lds_pos) {
        double *lp = reinterpret_cast<double *>(chain_dyn_lds);
#pragma unroll
        for (int d = 0; d < D; d++) lp[d * a.n + id] = p[d];
    } else {
#pragma unroll
        for (int d = 0; d < D; d++) a.pos[(size_t)d * a.n + id] = p[d];
    }
}

template <int D>
__device__ __forceinline__ void separation_vector(const ChainArgs &a, const double *ap, const double *tp, double *s) {
#pragma unroll
    for (int d = 0; d < D; d++) s[d] = py_mod((tp[d] - ap[d]) + a.half, a.L) - a.half;
}

// event_handler/abstracts/abstracts.py:82-95 for a velocity with one nonzero component
template <int D>
__device__ __forceinline__ void slice_pos(const ChainArgs &a, double *p, int dir, double speed, Time from, Time to) {
    double dt = time_sub(to, from);
#pragma unroll
    for (int d = 0; d < D; d++) {
        double v = d == dir ? speed : 0.0;
        p[d] = py_mod(p[d] + v * dt, a.L);
    }
}

__device__ __forceinline__ bool pot_draws(const ecmc_tagger_desc &tg) {
    if (tg.event_handler == ECMC_EH_TWO_LEAF_BOUNDING) return true;
    return tg.potential != ECMC_POT_HARD_SPHERE && tg.potential != ECMC_POT_HARD_DIPOLE;
}

// Bounding event rate of a cell-bounding candidate: the InnerPointEstimator bound of the
// target cell's offset from the active cell along dir, upper or lower by the sign of the
// charge correction factor (cell_bounding_potential.py:161-201).
template <int D>
__device__ __forceinline__ double cell_rate_of_cells(const ChainArgs &a, const ecmc_tagger_desc &tg, int acell, int tcell,
                                                     int dir, double ca, double ct) {
    int rel = 0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const int ti = (tcell / a.cum[d]) % a.cps[d], ai = (acell / a.cum[d]) % a.cps[d];
        rel += ((ti - ai) % a.cps[d] + a.cps[d]) % a.cps[d] * a.cum[d];
    }
    const double *row = a.cbt + (size_t)(rel * 3 + dir) * 2;
    if (tg.use_charge) {
        const double cf = ca * ct / tg.bpar[2];
        return cf > 0.0 ? row[0] * cf : row[1] * cf;
    }
    return row[0];
}
// The same for a target unit, its cell from its position.
template <int D>
__device__ __forceinline__ double cell_rate(const ChainArgs &a, const ecmc_tagger_desc &tg, int acell, int target, int dir,
                                            double ca, double ct) {
    double tp[D];
    load_pos<D>(a, target, tp);
    return cell_rate_of_cells<D>(a, tg, acell, pos_to_cell<D>(a, tp), dir, ca, ct);
}

// Is cell c among the nearby cells of cell ac (within `layers` per axis, periodic)?
__device__ __forceinline__ bool is_nearby(const ChainArgs &a, int dim, int c, int ac) {
    for (int d = 0; d < dim; d++) {
        const int ci = (c / a.cum[d]) % a.cps[d], ai = (ac / a.cum[d]) % a.cps[d];
        const int off = ((ci - ai) % a.cps[d] + a.cps[d]) % a.cps[d];
        if (off > a.layers && off < a.cps[d] - a.layers) return false;
    }
    return true;
}

// Kernel variants by the features of the systems they run (FM): bit p for the two-leaf pair
// potential ECMC_POT_p inverted exactly, FEAT_COULOMB for the merged-image Coulomb factors through
// the inverse-power Coulomb bound (thinning by the Ewald derivative), FEAT_FMAP for factor-type
// maps, FEAT_CELL_BOUNDING and FEAT_CELL_VETO for the two far-cell schemes.  A variant leaves the
// code of the other features out of the kernel (a smaller leg loop, fewer instruction-cache
// misses); the host launches the first variant that covers every feature of its systems.
constexpr int FEAT_COULOMB = 1 << 9, FEAT_FMAP = 1 << 10, FEAT_CELL_BOUNDING = 1 << 11, FEAT_CELL_VETO = 1 << 12;
constexpr int FEAT_ALL = (1 << 13) - 1;
constexpr int FEAT_EWALD = FEAT_COULOMB | FEAT_CELL_BOUNDING | FEAT_CELL_VETO;   // thinning by the Ewald derivative
constexpr int FEAT_IP_FILTER = (1 << ECMC_POT_INVERSE_POWER) | (1 << ECMC_POT_INVERSE_POWER_TRUNC);
constexpr int VARIANT_LJ = FEAT_IP_FILTER;
constexpr int VARIANT_CHARGED_CB = (1 << ECMC_POT_INVERSE_POWER) | FEAT_COULOMB | FEAT_CELL_BOUNDING;
constexpr int VARIANT_CHARGED_CV = (1 << ECMC_POT_INVERSE_POWER) | FEAT_COULOMB | FEAT_CELL_VETO;
constexpr int VARIANT_DIPOLES = (1 << ECMC_POT_INVERSE_POWER) | (1 << ECMC_POT_DISPLACED_EVEN_POWER) | FEAT_COULOMB | FEAT_FMAP;
#define POT_BUILT(p) ((FM >> (p)) & 1)

// Candidate time displacement of one pair (send_event_time of the pair event handlers).
template <int D, int FM>
__device__ __forceinline__ double pair_dt(const ChainArgs &a, const ecmc_tagger_desc &tg, int dir, double speed, const double *s,
                          double c1, double c2, double dU) {
    if (tg.event_handler == ECMC_EH_TWO_LEAF) {
        switch (tg.potential) {
        case ECMC_POT_HARD_SPHERE:
            if (!POT_BUILT(ECMC_POT_HARD_SPHERE)) return NAN;
            return hs_displacement<D>(dir, speed, s, tg.par[0]);
        case ECMC_POT_HARD_DIPOLE:
            if (!POT_BUILT(ECMC_POT_HARD_DIPOLE)) return NAN;
            return hd_displacement<D>(dir, speed, s, tg.par[0], tg.par[1]);
        case ECMC_POT_INVERSE_POWER:
            if (!POT_BUILT(ECMC_POT_INVERSE_POWER)) return NAN;
            return ip_displacement<D>(dir, speed, s, c1, c2, tg.par[0], tg.par[1], dU);
        case ECMC_POT_INVERSE_POWER_TRUNC:
            if (!POT_BUILT(ECMC_POT_INVERSE_POWER_TRUNC)) return NAN;
            return ipt_displacement<D>(dir, speed, s, c1, c2, tg.par[0], tg.par[1], tg.par[2], dU);
        case ECMC_POT_LENNARD_JONES: {
            if (!POT_BUILT(ECMC_POT_LENNARD_JONES)) return NAN;
            MexicanHat h = mh_lennard_jones(tg.par[0], tg.par[1]);
            return mh_displacement<D>(h, dir, speed, s, dU);
        }
        case ECMC_POT_DISPLACED_EVEN_POWER: {
            if (!POT_BUILT(ECMC_POT_DISPLACED_EVEN_POWER)) return NAN;
            MexicanHat h = mh_displaced_even_power(tg.par[0], tg.par[1], tg.par[2]);
            return mh_displacement<D>(h, dir, speed, s, dU);
        }
        default: return NAN;
        }
    }
    // two_leaf_unit_bounding_potential_event_handler.py:138-143 with the Coulomb bound
    if (!(FM & FEAT_COULOMB)) return NAN;
    double p[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; d++) p[d] = d < D ? s[(dir + d) % D] : 0.0;
    return ipcb_displacement(tg.bpar[0] * c1 * c2, p[0], p[1], p[2], dU, a.L) / speed;
}

// Candidate j of a factor-map tagger (factor_type_maps.py:325-356): a local factor's entries for
// the active leaf number k; a non-local factor's entries for every other composite object in
// identifier order, the entry's other leaf in that object (or, for numbers below P, in the
// active unit's own object).
__device__ __forceinline__ int fmap_target(const ecmc_tagger_desc &tg, int P, int act, int j) {
    const int m = act / P, k = act % P, fn = tg.fmap_n[k];
    if (tg.fmap_local) return m * P + tg.fmap_target[k][j];
    const int o = j / fn, e = j - o * fn;
    const int other = o < m ? o : o + 1;
    const int tl = tg.fmap_target[k][e];
    return tl < P ? m * P + tl : other * P + (tl - P);
}
__device__ __forceinline__ int fmap_count(const ecmc_tagger_desc &tg, int P, int n, int act) {
    const int fn = tg.fmap_n[act % P];
    return tg.fmap_local ? fn : (n / P - 1) * fn;
}

struct Best {
    double q, r;
    int32_t j;
};
__device__ __forceinline__ bool best_lt(const Best &x, const Best &y) {
    if (x.q < y.q) return true;
    if (x.q == y.q) {
        if (x.r < y.r) return true;
        if (x.r == y.r) return x.j < y.j;
    }
    return false;
}

// One butterfly step within the 16-lane rows by a DPP lane permutation (CTRL: 0xB1 swaps
// neighbours, 0x4E pairs of neighbours, 0x141 mirrors each half row, 0x140 the row): after the
// four steps every lane of a row holds the row's least (time, index) and its hash sum.
template <int CTRL>
__device__ __forceinline__ int dpp_move(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false); }
template <int CTRL>
__device__ __forceinline__ double dpp_move_d(double v) {
    return __hiloint2double(dpp_move<CTRL>(__double2hiint(v)), dpp_move<CTRL>(__double2loint(v)));
}
template <int CTRL>
__device__ __forceinline__ void best_dpp_step(Best &b, uint32_t &h) {
    Best o;
    o.q = dpp_move_d<CTRL>(b.q);
    o.r = dpp_move_d<CTRL>(b.r);
    o.j = dpp_move<CTRL>(b.j);
    const uint32_t oh = (uint32_t)dpp_move<CTRL>((int)h);
    if (best_lt(o, b)) b = o;
    h += oh;
}

// Collective workgroup min-reduce of (time, index) and sum of the hash terms: DPP steps within
// each row, shuffles across the rows, LDS across the wavefronts.  The order is total (indices
// are distinct) and the sum wraps, so the result does not depend on the reduction order.
__device__ __forceinline__ void block_reduce(Best &b, uint32_t &h, Best *sb, uint32_t *sh) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    best_dpp_step<0xB1>(b, h);
    best_dpp_step<0x4E>(b, h);
    best_dpp_step<0x141>(b, h);
    best_dpp_step<0x140>(b, h);
#pragma unroll
    for (int off = 32; off >= 16; off >>= 1) {
        Best o;
        o.q = __shfl_xor(b.q, off, 64);
        o.r = __shfl_xor(b.r, off, 64);
        o.j = __shfl_xor(b.j, off, 64);
        uint32_t oh = __shfl_xor(h, off, 64);
        if (best_lt(o, b)) b = o;
        h += oh;
    }
    if (lane == 0) { sb[wave] = b; sh[wave] = h; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / 64; w++) {
            if (best_lt(sb[w], sb[0])) sb[0] = sb[w];
            sh[0] += sh[w];
        }
    }
    __syncthreads();
    b = sb[0];
    h = sh[0];
}

// Collective workgroup minimum of a non-negative float (+inf allowed); sh holds NT / 64 words.
__device__ __forceinline__ float block_min_float(float v, uint32_t *sh) {
    v = fminf(v, __int_as_float(dpp_move<0xB1>(__float_as_int(v))));
    v = fminf(v, __int_as_float(dpp_move<0x4E>(__float_as_int(v))));
    v = fminf(v, __int_as_float(dpp_move<0x141>(__float_as_int(v))));
    v = fminf(v, __int_as_float(dpp_move<0x140>(__float_as_int(v))));
#pragma unroll
    for (int off = 32; off >= 16; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = __float_as_uint(v);
    __syncthreads();
    float r = __uint_as_float(sh[0]);
    for (int w = 1; w < NT / 64; w++) r = fminf(r, __uint_as_float(sh[w]));
    __syncthreads();
    return r;
}

// The candidate sets whose exact inversions are filtered by FP32 brackets first: two-leaf
// inverse-power factors (plain and truncated-and-shifted) of an integer power up to 32.  The
// Coulomb bound's sets are inverted exactly: its brackets (ipcb_time_bounds) cost nearly as much
// as its exact inversion, and filtering them measured slower.
__device__ __forceinline__ bool fp32_filter_applies(const ecmc_tagger_desc &tg) {
    if (tg.event_handler != ECMC_EH_TWO_LEAF) return false;
    if (tg.potential != ECMC_POT_INVERSE_POWER && tg.potential != ECMC_POT_INVERSE_POWER_TRUNC) return false;
    const double p = tg.par[0];
    return p >= 1.0 && p <= 32.0 && p == floor(p);
}

template <int D, int FM>
__device__ __forceinline__ void create_pair(const ChainArgs &a, ChainState &st, uint32_t *mt, const ecmc_tagger_desc &tg, int tix,
                            int *nb_cell, int *nb_pref, Best *sb, uint32_t *sh, int *scan) {
    const int tid = threadIdx.x;
    if (tg.kind == ECMC_TAG_PAIR_CELLS) {
        // the nearby cells (at most 27) and the prefix sums of their occupant counts: one lane
        // per cell of the first wavefront, an inclusive scan by shuffles
        if (tid < 64) {
            const int cell = st.occ_cell;
            const int cnt = a.nearby_n[cell];
            int cc = 0, num = 0;
            if (tid < cnt) {
                cc = a.nearby[(size_t)cell * a.nb_stride + tid];
                num = occn_base(a)[cc];
            }
            int incl = num;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o, 64);
                if (tid >= o) incl += v;
            }
            if (tid < cnt) {
                nb_cell[tid] = cc;
                nb_pref[tid + 1] = incl;
            }
            if (tid == 0) {
                nb_pref[0] = 0;
                nb_cell[31] = cnt;
            }
        }
    } else if (tid == 0) {
        nb_pref[0] = 0;
        nb_pref[1] = a.n;  // identifiers 0..n-1 minus the active one
        nb_cell[31] = 1;
    }
    __syncthreads();
    const int nc = nb_cell[31];
    const bool all = tg.kind == ECMC_TAG_PAIR_ALL;
    const bool sur = tg.kind == ECMC_TAG_SURPLUS;
    const bool far = (FM & FEAT_CELL_BOUNDING) && tg.kind == ECMC_TAG_CELL_BOUNDING;
    const bool fmap = (FM & FEAT_FMAP) && tg.kind == ECMC_TAG_FACTOR_MAP;
    // the far cells of this create (built and read within it): in LDS when the launch has room
    int32_t *const cand = a.lds_cand_off >= 0 ? reinterpret_cast<int32_t *>(chain_dyn_lds + a.lds_cand_off) : a.cand;
    if (far) {
        // cell_bounding_potential_tagger.py:150-155: the occupied cells that are not nearby the
        // active cell, in cell order; each threads scans a contiguous run of cells, an
        // exclusive scan of the counts places its candidates
        const int ac = st.occ_cell;
        const int chunk = (a.ncells + NT - 1) / NT;
        const int lo = min(tid * chunk, a.ncells), hi = min(lo + chunk, a.ncells);
        // the occupied far cells of this thread's run, as bits when the run has at most 32 cells
        const bool use_bits = chunk <= 32;
        uint32_t bits = 0;
        int cnt = 0;
        for (int c = lo; c < hi; c++)
            if (occn_base(a)[c] > 0 && !is_nearby(a, D, c, ac)) {
                cnt++;
                if (use_bits) bits |= 1u << (c - lo);
            }
        // exclusive scan of the counts over the workgroup: shuffles within each wavefront, the
        // wavefront totals in scan[NT + 1 ..]
        const int lane = tid & 63, wave = tid >> 6;
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        if (lane == 63) scan[NT + 1 + wave] = incl;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; w++) before += scan[NT + 1 + w];
        int k = before + incl - cnt;
        if (tid == NT - 1) scan[NT] = before + incl;
        bool multi = false;
        for (int c = lo; c < hi; c++)
            if (use_bits ? ((bits >> (c - lo)) & 1u) != 0 : (occn_base(a)[c] > 0 && !is_nearby(a, D, c, ac))) {
                if (occn_base(a)[c] != 1) multi = true;
                cand[k++] = c;   // the cell; its one occupant is the candidate
            }
        if (multi) st.err = ECMC_ERR_STATE;   // the handler takes exactly one target unit
        __syncthreads();
    }
    const int M = fmap ? (st.active < 0 ? 0 : fmap_count(tg, a.leaves, a.n, st.active))
                       : (all ? a.n - 1 : (sur ? st.nsur : (far ? scan[NT] : nb_pref[nc])));
    const bool draws = pot_draws(tg);
    ring_ensure(st, mt, a, draws ? 2ull * (uint64_t)M : 0ull);
    const uint64_t base = st.cursor;
    const int act = st.active, dir = st.adir;
    const double speed = st.aspeed;
    const Time at = st.at;
    double ap[D];
    load_pos<D>(a, act, ap);
    const double ca = tg.use_charge ? charge_of(a, act) : 1.0;
    const int jbase = st.ncand;
    Best best = {INFINITY, INFINITY, 0x7fffffff};
    uint32_t hsum = 0;
    auto cand_target = [&](int j) {
        if (fmap) return fmap_target(tg, a.leaves, act, j);
        if (all) return j < act ? j : j + 1;
        if (sur) return (int)a.sur_id[j];
        if (far) return (int)occ_base(a)[(size_t)cand[j] * a.occ_cap];
        // the nearby cell holding candidate j: the largest c with nb_pref[c] <= j (binary search
        // over the nc + 1 prefix sums; empty cells repeat a prefix)
        int lo = 0, hi = nc;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (nb_pref[mid] <= j) lo = mid;
            else hi = mid;
        }
        return (int)occ_base(a)[(size_t)nb_cell[lo] * a.occ_cap + (j - nb_pref[lo])];
    };
    // FP32 filter: every candidate's time displacement is bracketed [lb, ub] from its separation
    // and the leading word of its draw; a candidate whose lb exceeds the smallest ub of the set
    // by more than the rounding of the active time can make up ends strictly later, so it can
    // be neither the earliest nor tie with it and is not inverted exactly; lb = +inf marks a
    // candidate without an event.  The earliest candidate and every candidate of equal time are
    // inverted exactly as before, so the result is the unfiltered one, unless a candidate left
    // out would have raised the reference's ValueError (a negative square root from rounding),
    // which the exact inversion flags.
    const bool filt = (FM & FEAT_IP_FILTER) && fp32_filter_applies(tg) && M > 1;
    float lbk0 = 0.f, lbk1 = 0.f;
    double cull_above = INFINITY;
    const int fp = (int)tg.par[0];
    const float two_over_p = 2.f / (float)fp;
    const bool ftrunc = tg.potential == ECMC_POT_INVERSE_POWER_TRUNC;
    const float rc2f = ftrunc ? (float)(tg.par[2] * tg.par[2]) : 0.f;
    const float ibf = (float)(1.0 / a.beta), isf = (float)(1.0 / speed);
    // the separation of the brackets: the difference wrapped by one box length, within two ulps
    // of L of separation_vector's Python-modulo form except near +-L/2, where the candidate keeps
    // the trivial brackets
    const double edge = a.half * (1.0 - 0x1p-40);
    const float sep_slack = (float)(a.L * 0x1p-49);
    auto cand_bounds = [&](int j, int target, float &lb, float &ub) {
        double tp[D], s[D];
        load_pos<D>(a, target, tp);
        bool near_edge = false;
#pragma unroll
        for (int d = 0; d < D; d++) {
            double v = tp[d] - ap[d];
            v = v >= a.half ? v - a.L : (v < -a.half ? v + a.L : v);
            near_edge = near_edge || !(fabs(v) < edge);
            s[d] = v;
        }
        const double ct = tg.use_charge ? charge_of(a, target) : 1.0;
        lb = 0.f;
        ub = INFINITY;
        const double pp = tg.par[1] * (ca * ct);
        const float K = (float)fabs(pp);
        if (!near_edge && K >= 0x1p-100f && K <= 0x1p100f) {
            const uint32_t w0 = ring_word(a, base + 2ull * j);
            ip_time_bounds<D>(s, dir, K, pp > 0.0, fp, two_over_p, ftrunc, dn(rc2f), up(rc2f),
                              dU_lower(w0, ibf * (1.f - 0x1p-22f)), dU_upper(w0, ibf * (1.f + 0x1p-22f)),
                              isf * (1.f - 0x1p-22f), isf * (1.f + 0x1p-22f), lb, ub, sep_slack);
        }
    };
    if (filt) {
        float ubmin = INFINITY;
        // two candidates per iteration, evaluated side by side (the second one a repeat of the
        // first past the end of the set, its result dropped)
        for (int j0 = tid; j0 < M; j0 += 2 * NT) {
            const int j1 = j0 + NT;
            const bool has1 = j1 < M;
            const int t0 = cand_target(j0), t1 = cand_target(has1 ? j1 : j0);
            float lb0, ub0, lb1, ub1;
            cand_bounds(j0, t0, lb0, ub0);
            cand_bounds(has1 ? j1 : j0, t1, lb1, ub1);
            if (j0 == tid) {
                lbk0 = lb0;
                lbk1 = lb1;
            }
            ubmin = fminf(ubmin, ub0);
            hsum += cand_mix((uint32_t)(jbase + j0), (uint32_t)t0);
            if (has1) {
                ubmin = fminf(ubmin, ub1);
                hsum += cand_mix((uint32_t)(jbase + j1), (uint32_t)t1);
            }
        }
        ubmin = block_min_float(ubmin, sh);
        // a displacement beyond ubmin + 2^-48 (1 + ubmin) ends later than ubmin does: the active
        // time's remainder lies in [0, 1), so the two sums with it round to doubles at least one
        // ulp apart, and Time's divmod by one keeps their order strict
        if (ubmin < INFINITY) cull_above = (double)ubmin + 0x1p-48 * (1.0 + (double)ubmin);
    }
    int kk = 0;
    for (int j = tid; j < M; j += NT, kk++) {
        if (filt) {
            float lb, ub;
            if (kk == 0) lb = lbk0;
            else if (kk == 1) lb = lbk1;
            else cand_bounds(j, cand_target(j), lb, ub);
            if (lb == INFINITY) {
                const Best c = {INFINITY, INFINITY, j};
                if (best_lt(c, best)) best = c;
                continue;
            }
            // lb = 0 never culls: the exact FP64 displacement can round below zero there
            if (lb > 0.f && (double)lb > cull_above) continue;
        }
        const int target = cand_target(j);
        double tp[D], s[D];
        load_pos<D>(a, target, tp);
        separation_vector<D>(a, ap, tp, s);
        double dU = 0.0;
        if (draws) {
            double u = words_to_random(ring_word(a, base + 2ull * j), ring_word(a, base + 2ull * j + 1));
            dU = -cr_log(1.0 - u) / a.beta;
        }
        const double ct = tg.use_charge ? charge_of(a, target) : 1.0;
        double dt;
        if ((FM & FEAT_CELL_BOUNDING) && tg.event_handler == ECMC_EH_CELL_BOUNDING) {
            // two_leaf_unit_cell_bounding_potential_event_handler.py:171-177
            const double rate = far ? cell_rate_of_cells<D>(a, tg, st.occ_cell, (int)cand[j], dir, ca, ct)
                                    : cell_rate<D>(a, tg, st.occ_cell, target, dir, ca, ct);
            dt = (rate > 0 ? dU / rate : INFINITY) / speed;
        } else {
            dt = pair_dt<D, FM>(a, tg, dir, speed, s, ca, ct, dU);
        }
        Time t = time_add(at, dt);
        Best c = {t.q, t.r, j};
        if (isnan(dt)) c.q = -INFINITY;  // flags a ValueError of the reference code
        if (best_lt(c, best)) best = c;
        if (!filt) hsum += cand_mix((uint32_t)(jbase + j), (uint32_t)target);
    }
    block_reduce(best, hsum, sb, sh);
    if (tid == 0) {
        if (draws) st.cursor += 2ull * (uint64_t)M;
        st.ncand += M;
        st.chash += hsum;
        Slot &sl = st.slot[tix];
        if (M == 0) {
            sl.valid = 0;
        } else if (best.q == -INFINITY) {
            st.err = ECMC_ERR_NUMERIC;
        } else {
            const int j = best.j, target = cand_target(j);
            sl.valid = 1;
            sl.cb_cell = st.occ_cell;
            sl.t.q = best.q;
            sl.t.r = best.r;
            sl.seq = st.seq + j;
            sl.target = target;
            sl.adir = dir;
            sl.aspeed = speed;
            double tp[D];
            load_pos<D>(a, target, tp);
            double p[D];
#pragma unroll
            for (int d = 0; d < D; d++) { p[d] = ap[d]; sl.tpos[d] = tp[d]; }
            if (!isinf(sl.t.q)) slice_pos<D>(a, p, dir, speed, at, sl.t);
#pragma unroll
            for (int d = 0; d < D; d++) sl.apos[d] = p[d];
            sl.astamp = sl.t;
        }
        st.seq += M;
    }
    __syncthreads();
}

// Collective: moves the entries [from, to) of the parallel arrays x and y (y may be NULL) by
// one slot up (by > 0) or down, in chunks of one entry per thread, every chunk read before it
// is written and the chunks ordered so that no entry is overwritten before it was read.
__device__ __forceinline__ void shift_entries(int32_t *x, int32_t *y, int from, int to, int by) {
    const int tid = threadIdx.x;
    if (by > 0) {
        for (int top = to; top > from; top -= NT) {
            const int lo = top - NT > from ? top - NT : from, i = lo + tid;
            const bool act = i < top;
            int32_t vx = 0, vy = 0;
            if (act) { vx = x[i]; if (y) vy = y[i]; }
            __syncthreads();
            if (act) { x[i + 1] = vx; if (y) y[i + 1] = vy; }
            __syncthreads();
        }
    } else {
        for (int lo = from; lo < to; lo += NT) {
            const int i = lo + tid;
            const bool act = i < to;
            int32_t vx = 0, vy = 0;
            if (act) { vx = x[i]; if (y) vy = y[i]; }
            __syncthreads();
            if (act) { x[i - 1] = vx; if (y) y[i - 1] = vy; }
            __syncthreads();
        }
    }
}

// Collective: the largest (want_max) or smallest index i < count with x[i] == vx and, when y is
// given, y[i] == vy; -1 when there is none.  red holds NT / 64 ints of workgroup scratch.
__device__ __forceinline__ int find_entry(const int32_t *x, const int32_t *y, int count, int32_t vx, int32_t vy, bool want_max,
                          int *red) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int best = want_max ? -1 : 0x7fffffff;
    for (int i = tid; i < count; i += NT)
        if (x[i] == vx && (!y || y[i] == vy)) best = want_max ? (i > best ? i : best) : (i < best ? i : best);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int v = __shfl_xor(best, o, 64);
        best = want_max ? (v > best ? v : best) : (v < best ? v : best);
    }
    if (lane == 0) red[wave] = best;
    __syncthreads();
    int r = red[0];
    for (int w = 1; w < NT / 64; w++) r = want_max ? (red[w] > r ? red[w] : r) : (red[w] < r ? red[w] : r);
    __syncthreads();
    return r == 0x7fffffff ? -1 : r;
}

// Collective, single_active_cell_occupancy.py:149-203: the unit that stops being active joins
// its cell's occupants, or the surplus when the cell is full (self._surplus.setdefault(cell,
// []).append(id) on the flat list in dict order: after the cell's last surplus entry, or at the
// end); the new active unit leaves the occupants or the surplus (the surplus never moves into
// the occupants: that branch of the reference needs an empty surplus list, which it deletes
// right away).  The list searches and shifts run over the workgroup.
template <int D>
__device__ __forceinline__ void occupancy_update(const ChainArgs &a, ChainState &st, int *red) {
    const int tid = threadIdx.x;
    const int id = st.active, old_id = st.occ_id, old_cell = st.occ_cell;
    const int n_old = (id != old_id && old_id >= 0) ? occn_base(a)[old_cell] : 0;
    double p[D];
    load_pos<D>(a, id, p);
    const int c = pos_to_cell<D>(a, p);
    __syncthreads();   // every thread has read the state and the counts before thread 0 changes them
    if (id == old_id) {
        if (tid == 0) {
            if (c < 0) st.err = ECMC_ERR_STATE;
            else st.occ_cell = c;
        }
        __syncthreads();
        return;
    }
    if (old_id >= 0) {
        const int n = n_old;
        if (a.max_occ <= 0 || n < a.max_occ) {
            if (tid == 0) {
                if (n >= a.occ_cap) st.err = ECMC_ERR_STATE;
                else {
                    occ_base(a)[(size_t)old_cell * a.occ_cap + n] = old_id;
                    occn_base(a)[old_cell] = n + 1;
                }
            }
        } else {
            const int nsur = st.nsur;
            if (nsur >= a.sur_cap) {
                if (tid == 0) st.err = ECMC_ERR_STATE;
            } else {
                const int last = find_entry(a.sur_cell, nullptr, nsur, old_cell, 0, true, red);
                const int at = last >= 0 ? last + 1 : nsur;
                shift_entries(a.sur_cell, a.sur_id, at, nsur, +1);
                if (tid == 0) {
                    a.sur_cell[at] = old_cell;
                    a.sur_id[at] = old_id;
                    st.nsur = nsur + 1;
                }
            }
        }
        __syncthreads();
        if (st.err) return;
    }
    if (c < 0) {
        if (tid == 0) st.err = ECMC_ERR_STATE;
        __syncthreads();
        return;
    }
    int32_t *Lc = occ_base(a) + (size_t)c * a.occ_cap;
    const int n = occn_base(a)[c];
    const int i = find_entry(Lc, nullptr, n, id, 0, false, red);
    if (i >= 0) {
        shift_entries(Lc, nullptr, i + 1, n, -1);
        if (tid == 0) occn_base(a)[c] = n - 1;
    } else {
        const int nsur = st.nsur;
        const int k = find_entry(a.sur_cell, a.sur_id, nsur, c, id, false, red);
        if (k < 0) {
            if (tid == 0) st.err = ECMC_ERR_STATE;
        } else {
            shift_entries(a.sur_cell, a.sur_id, k + 1, nsur, -1);
            if (tid == 0) st.nsur = nsur - 1;
        }
    }
    if (tid == 0) {
        st.occ_cell = c;
        st.occ_id = id;
    }
    __syncthreads();
}

// Creates of one non-pair tagger (thread 0 only, ring words made available by the caller).
template <int D>
__device__ __forceinline__ void create_single(const ChainArgs &a, ChainState &st, const ecmc_tagger_desc &tg, int tix) {
    Slot &sl = st.slot[tix];
    switch (tg.kind) {
    case ECMC_TAG_CELL_BOUNDARY: {
        // cell_boundary_event_handler.py:98-156
        double p[D];
        load_pos<D>(a, st.active, p);
        int cell = pos_to_cell<D>(a, p);
        if (cell < 0) { st.err = ECMC_ERR_STATE; return; }
        int dir = st.adir;
        int id = (cell / a.cum[dir]) % a.cps[dir];
        int nid = (id + 1) % a.cps[dir];
        int nb = cell + (nid - id) * a.cum[dir];
        double boundary = a.cmin[(size_t)nb * D + dir];
        double sep = boundary - p[dir];
        if (sep < 0.0) sep = sep + a.L;
        double dt = sep / st.aspeed;
        sl.t = time_add(st.at, dt);
#pragma unroll
        for (int d = 0; d < D; d++) sl.apos[d] = p[d];
        sl.astamp = st.at;
        sl.cb_boundary = boundary;
        sl.cb_dir = dir;
        sl.adir = dir;
        sl.aspeed = st.aspeed;
        sl.valid = 1;
        sl.seq = st.seq++;
        break;
    }
    case ECMC_TAG_SAMPLING:
        st.samp_time = time_add(st.samp_time, tg.value);
        sl.t = st.samp_time;
        sl.valid = 1;
        sl.seq = st.seq++;
        break;
    case ECMC_TAG_END_OF_RUN:
        sl.t = time_from_float(tg.value);
        sl.valid = 1;
        sl.seq = st.seq++;
        break;
    case ECMC_TAG_START_OF_RUN:
        sl.t.q = 0.0;
        sl.t.r = 0.0;
        sl.valid = 1;
        sl.seq = st.seq++;
        break;
    default: break;
    }
}

// end of chain: chain time and the randint of the next active unit
// (single_independent_active_periodic_direction_end_of_chain_event_handler.py:176-237): with
// composite objects the root's randint, then the leaf's.  Returns false when the ring ran out of
// words before the draws completed; the root already drawn is kept in the slot (cv_index).
__device__ __forceinline__ bool create_end_of_chain(const ChainArgs &a, ChainState &st, const ecmc_tagger_desc &tg, int tix) {
    Slot &sl = st.slot[tix];
    if (!sl.valid || sl.target >= 0x40000000) {
        if (!sl.valid || sl.target != 0x40000000) {
            double since = time_sub(st.at, st.eoc_last);
            if (!(0.0 <= since && since <= tg.value)) { st.err = ECMC_ERR_CONFIGURATION; return true; }
            double chain = time_sub(st.eoc_last, st.at) + tg.value;
            sl.t = time_add(st.at, chain);
            sl.valid = 1;
            sl.target = 0x40000000;  // draws pending
            sl.cv_index = -1;        // no root drawn yet
        }
    }
    const int P = a.leaves;
    while (true) {
        const bool root_phase = sl.cv_index < 0;
        const uint32_t nrange = root_phase ? (uint32_t)(a.n / P) : (uint32_t)P;
        int k = 0;
        while ((nrange >> k) != 0u) k++;
        bool got = false;
        uint32_t v = 0;
        while (st.cursor < st.produced) {
            uint32_t w = ring_word(a, st.cursor);
            st.cursor++;
            v = k >= 32 ? w : (w >> (32 - k));
            if (v < nrange) { got = true; break; }
        }
        if (!got) return false;
        if (root_phase && P > 1) {
            sl.cv_index = (int32_t)v;
            continue;
        }
        sl.target = root_phase ? (int32_t)v : sl.cv_index * P + (int32_t)v;
        sl.seq = st.seq++;
        return true;
    }
}

// cell veto send_event_time (cell_veto_event_handler.py:190-238): the Walker table by the sign
// of the charge correction factor; random.choice over its entries (rejection on bit_length
// words), uniform(0, mean) for item or alias, then expovariate / (total rate * speed).
// Thread 0 only.  Returns false when the ring ran out of words; the drawn entry is kept in the
// slot, so the call resumes after a refill.
template <int D>
__device__ __forceinline__ bool create_cell_veto(const ChainArgs &a, ChainState &st, const ecmc_tagger_desc &tg, int tix) {
    Slot &sl = st.slot[tix];
    const int act = st.active, dir = st.adir;
    double cf = (tg.use_charge ? charge_of(a, act) : 1.0) * 1.0 / tg.bpar[2];
    int b = 0;
    if (!(cf > 0.0)) { cf *= -1.0; b = 1; }
    const int wt = dir * 2 + b;
    const uint32_t m = (uint32_t)a.cv_m;
    if (sl.cv_phase != 2) {
        sl.cv_phase = 1;
        int k = 0;
        while ((m >> k) != 0u) k++;
        while (st.cursor < st.produced) {
            const uint32_t w = ring_word(a, st.cursor);
            st.cursor++;
            const uint32_t v = k >= 32 ? w : (w >> (32 - k));
            if (v < m) { sl.cv_index = (int32_t)v; sl.cv_phase = 2; break; }
        }
        if (sl.cv_phase != 2) return false;
    }
    if (st.produced - st.cursor < 4) return false;
    const CvEntry e = a.cvw[(size_t)wt * m + sl.cv_index];
    const double u = 0.0 + (a.cv_mean[wt] - 0.0) * words_to_random(ring_word(a, st.cursor), ring_word(a, st.cursor + 1));
    const int rel = u <= e.rate0 ? e.item0 : e.item1;
    const double u2 = words_to_random(ring_word(a, st.cursor + 2), ring_word(a, st.cursor + 3));
    const double dU = -cr_log(1.0 - u2) / a.beta;
    st.cursor += 4;
    sl.cv_phase = 0;
    const double total = a.cv_total[wt] * cf;
    const double row = a.cbt[(size_t)(rel * 3 + dir) * 2 + b];
    sl.cv_rate = (b ? -row : row) * cf;
    if (!(sl.cv_rate > 0.0) || rel < 0) { st.err = ECMC_ERR_NUMERIC; return true; }
    double ap[D];
    load_pos<D>(a, act, ap);
    const int acell = pos_to_cell<D>(a, ap);
    // cuboid_periodic_cells.py:182-207: the cell centre shifted by the relative cell's lower
    // corner lies half a cell from every wall, so the per-axis identifier sum is that cell
    int tcell = 0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const int ai = (acell / a.cum[d]) % a.cps[d], ri = (rel / a.cum[d]) % a.cps[d];
        tcell += ((ai + ri) % a.cps[d]) * a.cum[d];
    }
    const double dt = dU / (total * st.aspeed);
    sl.t = time_add(st.at, dt);
    sl.valid = 1;
    sl.seq = st.seq++;
    sl.target = -1;
    sl.cb_cell = tcell;
    sl.adir = dir;
    sl.aspeed = st.aspeed;
    if (!isinf(sl.t.q)) slice_pos<D>(a, ap, dir, st.aspeed, st.at, sl.t);
#pragma unroll
    for (int d = 0; d < D; d++) sl.apos[d] = ap[d];
    sl.astamp = sl.t;
    return true;
}

// ewald_derivative over one wavefront, all 64 lanes calling with the same separation.  The
// host lists the terms of the sequential loop (the real-space images by k, j, i, then the
// Fourier terms by i, j, k); term g belongs to lane g % 64, slot g / 64.  Each lane evaluates
// its terms with the sequential loop's operations (a Fourier factor by stepping the same
// cos/sin recurrence from its start as far as that term), adds its slots in order, and a
// butterfly over the lanes adds the 64 partial sums (commutative pairings, so every lane ends
// with the same sum).  The order of the additions differs from the sequential loop's, so the
// result agrees with it to rounding.  More than 256 terms fall back to the sequential function.
// The sequential form where the wave form does not apply: called, not inlined (a rare path).
__device__ __noinline__ static double ewald_derivative_call(const Ewald &e, double sx, double sy, double sz) {
    return ewald_derivative(e, sx, sy, sz);
}
__device__ __forceinline__ double ewald_derivative_wave(const Ewald &e, const int32_t *codes, int ncodes, double sx, double sy,
                                        double sz) {
    if (ncodes > 256 || !codes) return ewald_derivative_call(e, sx, sy, sz);
    const int lane = threadIdx.x & 63;
    const int n = e.fc + 1;
    const double L = e.L;
    int code[4];
#pragma unroll
    for (int s = 0; s < 4; s++) code[s] = lane + 64 * s < ncodes ? codes[lane + 64 * s] : -1;
    // cos and sin of tpol s_d on lane d (the same functions per component), broadcast
    double cd, sd;
    sincos(e.tpol * (lane == 0 ? sx : (lane == 1 ? sy : sz)), &sd, &cd);
    const double dcx = __shfl(cd, 0, 64), dsx = __shfl(sd, 0, 64);
    const double dcy = __shfl(cd, 1, 64), dsy = __shfl(sd, 1, 64);
    const double dcz = __shfl(cd, 2, 64), dsz = __shfl(sd, 2, 64);
    double sum = 0.0;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        double v = 0.0;
        const int c = code[s];
        if (c >= 0) {
            const int i = ((c >> 16) & 255) - 128, j = ((c >> 8) & 255) - 128, k = (c & 255) - 128;
            if (c >> 24) {
                double tmp, cxv = dcx, sxv = dsx, cyv = 1.0, syv = 0.0, czv = 1.0, szv = 0.0;
                for (int r = 1; r < i; r++) { tmp = cxv; cxv = tmp * dcx - sxv * dsx; sxv = sxv * dcx + tmp * dsx; }
                for (int r = 0; r < j; r++) { tmp = cyv; cyv = tmp * dcy - syv * dsy; syv = syv * dcy + tmp * dsy; }
                for (int r = 0; r < k; r++) { tmp = czv; czv = tmp * dcz - szv * dsz; szv = szv * dcz + tmp * dsz; }
                v = e.f[(i * n + j) * n + k] * sxv * cyv * czv;
            } else {
                const double vz2 = (sz + k * L) * (sz + k * L);
                const double vy2 = (sy + j * L) * (sy + j * L);
                const double vx = sx + i * L;
                const double v2 = vx * vx + vy2 + vz2;
                const double vn = sqrt(v2);
                v = vx * (e.two_aol_rpi * exp(-e.aol2 * v2) + erfc(e.aol * vn) / vn) / v2;
            }
        }
        sum += v;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    return sum;
}

// The terms of ewald_derivative in its loop order, as the codes of ewald_derivative_wave.
static std::vector<int32_t> ewald_term_codes(int fc, int pc) {
    std::vector<int32_t> out;
    for (int k = -pc; k < pc + 1; k++) {
        const int cy = (int)sqrt((double)(pc * pc - k * k));
        for (int j = -cy; j < cy + 1; j++) {
            const int cx = (int)sqrt((double)(pc * pc - j * j - k * k));
            for (int i = -cx; i < cx + 1; i++) out.push_back(((i + 128) << 16) | ((j + 128) << 8) | (k + 128));
        }
    }
    for (int i = 1; i < fc + 1; i++) {
        const int cy = (int)sqrt((double)(fc * fc - i * i));
        for (int j = 0; j < cy + 1; j++) {
            const int cx = (int)sqrt((double)(fc * fc - i * i - j * j));
            for (int k = 0; k < cx + 1; k++) out.push_back((1 << 24) | ((i + 128) << 16) | ((j + 128) << 8) | (k + 128));
        }
    }
    return out;
}

// The separation (permuted: x along the motion) at which the commit of tagger w's stored
// event evaluates the merged-image Coulomb derivative, or false when it evaluates none.
template <int D>
__device__ __forceinline__ bool thinning_separation(const ChainArgs &a, const ChainState &st, const ecmc_tagger_desc &T, int w,
                                    double *q3) {
    const Slot &sl = st.slot[w];
    double s[D];
    if ((T.kind == ECMC_TAG_PAIR_CELLS || T.kind == ECMC_TAG_PAIR_ALL || T.kind == ECMC_TAG_SURPLUS ||
         T.kind == ECMC_TAG_CELL_BOUNDING || T.kind == ECMC_TAG_FACTOR_MAP) &&
        (T.event_handler == ECMC_EH_CELL_BOUNDING || T.event_handler == ECMC_EH_TWO_LEAF_BOUNDING)) {
        separation_vector<D>(a, sl.apos, sl.tpos, s);
    } else if (T.kind == ECMC_TAG_CELL_VETO) {
        if (occn_base(a)[sl.cb_cell] != 1) return false;
        double tp[D];
        load_pos<D>(a, occ_base(a)[(size_t)sl.cb_cell * a.occ_cap], tp);
        separation_vector<D>(a, sl.apos, tp, s);
    } else {
        return false;
    }
    q3[0] = q3[1] = q3[2] = 0.0;
    for (int d = 0; d < D; d++) q3[d] = s[(sl.adir + d) % D];
    return true;
}

// Shared memory of the leg loop's collective steps.
struct LegShared {
    int nb_cell[32];
    int nb_pref[33];
    Best sb[NT / 64];
    uint32_t sh[NT / 64];
    int scan[NT + 1 + NT / 64];
    int flag;
};

// The event handlers of tagger t send their candidate event times: the pair taggers' candidate
// sets over the workgroup, the cell veto and the end of chain with their bounded ring refills,
// the others on thread 0 (collective; tag_activator.py:228-289 for one tag of a create list).
template <int D, int FM>
__device__ __forceinline__ void create_tagger(const ChainArgs &a, ChainState &st, uint32_t *mt, const ecmc_tagger_desc &c, int t,
                              LegShared &ls) {
    const int tid = threadIdx.x;
    if (c.kind == ECMC_TAG_PAIR_CELLS || c.kind == ECMC_TAG_PAIR_ALL || c.kind == ECMC_TAG_SURPLUS ||
        c.kind == ECMC_TAG_CELL_BOUNDING || c.kind == ECMC_TAG_FACTOR_MAP) {
        create_pair<D, FM>(a, st, mt, c, t, ls.nb_cell, ls.nb_pref, ls.sb, ls.sh, ls.scan);
    } else if ((FM & FEAT_CELL_VETO) && c.kind == ECMC_TAG_CELL_VETO) {
        if (tid == 0) { st.slot[t].valid = 0; st.slot[t].cv_phase = 0; }
        __syncthreads();
        // the same bounded refill loop as the end-of-chain randint below
        for (int attempt = 0;; attempt++) {
            ring_ensure(st, mt, a, 64);
            if (tid == 0) {
                ls.flag = (st.active < 0 || create_cell_veto<D>(a, st, c, t)) ? 1 : 0;
                if (!ls.flag && attempt >= 64) { st.err = ECMC_ERR_STATE; ls.flag = 1; }
            }
            __syncthreads();
            if (ls.flag) break;
            ring_ensure(st, mt, a, 624);
        }
    } else if (c.kind == ECMC_TAG_END_OF_CHAIN) {
        if (tid == 0) st.slot[t].valid = 0;
        __syncthreads();
        // randint's rejection loop: each word is accepted with probability >= 1/2,
        // so 64 refills (40k words) without success only happen after a defect
        for (int attempt = 0;; attempt++) {
            ring_ensure(st, mt, a, 64);
            if (tid == 0) {
                ls.flag = create_end_of_chain(a, st, c, t) ? 1 : 0;
                if (!ls.flag && attempt >= 64) { st.err = ECMC_ERR_STATE; ls.flag = 1; }
            }
            __syncthreads();
            if (ls.flag) break;
            ring_ensure(st, mt, a, 624);
        }
    } else {
        if (tid == 0) create_single<D>(a, st, c, t);
        __syncthreads();
    }
}

__device__ __forceinline__ bool commit_thins(const ecmc_tagger_desc &T) {
    return ((T.kind == ECMC_TAG_PAIR_CELLS || T.kind == ECMC_TAG_PAIR_ALL || T.kind == ECMC_TAG_SURPLUS ||
             T.kind == ECMC_TAG_CELL_BOUNDING || T.kind == ECMC_TAG_FACTOR_MAP) &&
            (T.event_handler == ECMC_EH_TWO_LEAF_BOUNDING || T.event_handler == ECMC_EH_CELL_BOUNDING)) ||
           T.kind == ECMC_TAG_CELL_VETO;
}

// The out-state of tagger w's stored event, committed to the state (thread 0; the caller made
// two ring words available when the event thins): send_out_state of its event handler and
// insert_into_global_state (mediator/single_process_mediator.py:130-150).
template <int D, int FM>
__device__ __forceinline__ void commit_event(const ChainArgs &a, ChainState &st, const ecmc_tagger_desc &T, int w, ecmc_event &ev,
                             double ewald_pre) {
    Slot &sl = st.slot[w];
    ev.tagger = w;
    ev.flags = 0;
    ev.active = st.active;
    ev.target = -1;
    ev.tq = sl.t.q;
    ev.tr = sl.t.r;
    ev.ncand = st.ncand;
    ev.chash = st.chash;
    const int act = st.active;
    double p[D];
    switch (T.kind) {
    case ECMC_TAG_PAIR_CELLS:
    case ECMC_TAG_PAIR_ALL:
    case ECMC_TAG_SURPLUS:
    case ECMC_TAG_CELL_BOUNDING:
    case ECMC_TAG_FACTOR_MAP: {
        ev.target = sl.target;
        bool lift = true;
        double s[D];
        separation_vector<D>(a, sl.apos, sl.tpos, s);
        if ((FM & FEAT_CELL_BOUNDING) && T.event_handler == ECMC_EH_CELL_BOUNDING) {
            // two_leaf_unit_cell_bounding_potential_event_handler.py:181-211: the active
            // unit is still in its cell; bound = the candidate's rate; thinning as below
            if (pos_to_cell<D>(a, sl.apos) != sl.cb_cell) { st.err = ECMC_ERR_STATE; break; }
            double q3[3] = {0, 0, 0};
            for (int d = 0; d < D; d++) q3[d] = s[(sl.adir + d) % D];
            const double c1 = T.use_charge ? charge_of(a, act) : 1.0;
            const double c2 = T.use_charge ? charge_of(a, sl.target) : 1.0;
            double bound = cell_rate<D>(a, T, sl.cb_cell, sl.target, sl.adir, c1, c2) * sl.aspeed;
            const double der = isnan(ewald_pre) ? ewald_derivative_call(a.ewald, q3[0], q3[1], q3[2]) : ewald_pre;
            double real = T.par[3] * c1 * c2 * der * sl.aspeed;
            lift = false;
            if (real > 0) {
                double u = words_to_random(ring_word(a, st.cursor), ring_word(a, st.cursor + 1));
                st.cursor += 2;
                u = 0 + (bound - 0) * u;
                if (u < real) lift = true;
            }
            if (!lift) st.est[6] += 1.0;
        } else if ((FM & FEAT_COULOMB) && T.event_handler == ECMC_EH_TWO_LEAF_BOUNDING) {
            // event_handler_with_bounding_potential.py:75-101
            double q3[3] = {0, 0, 0};
            for (int d = 0; d < D; d++) q3[d] = s[(sl.adir + d) % D];
            const double c1 = T.use_charge ? charge_of(a, act) : 1.0;
            const double c2 = T.use_charge ? charge_of(a, sl.target) : 1.0;
            double bound = ipcb_derivative(T.bpar[0] * c1 * c2, q3[0], q3[1], q3[2]) * sl.aspeed;
            const double der = isnan(ewald_pre) ? ewald_derivative_call(a.ewald, q3[0], q3[1], q3[2]) : ewald_pre;
            double real = T.par[3] * c1 * c2 * der * sl.aspeed;
            lift = false;
            if (real > 0) {
                double u = words_to_random(ring_word(a, st.cursor), ring_word(a, st.cursor + 1));
                st.cursor += 2;
                u = 0 + (bound - 0) * u;
                if (u < real) lift = true;
            }
            if (!lift) st.est[6] += 1.0;
        }
        store_pos<D>(a, act, sl.apos);
        st.at = sl.astamp;
        if (lift) {
            st.est[0] += s[sl.adir];
            st.est[4] += 1.0;
            st.active = sl.target;
            ev.flags |= 1;
        }
        st.est[2] += 1.0;
        break;
    }
    case ECMC_TAG_CELL_VETO: {
        if (!(FM & FEAT_CELL_VETO)) { st.err = ECMC_ERR_CONFIGURATION; break; }
        // mediator.py:265-292: the occupant of the sampled cell, or none;
        // leaf_unit_cell_veto_event_handler.py:127-149 with
        // event_handler_with_bounding_potential.py:75-101: thinning of the stored rate
        const int tc = sl.cb_cell;
        const int no = occn_base(a)[tc];
        if (no > 1) { st.err = ECMC_ERR_STATE; break; }
        bool lift = false;
        double s[D];
#pragma unroll
        for (int d = 0; d < D; d++) s[d] = 0.0;
        if (no == 1) {
            const int tid2 = occ_base(a)[(size_t)tc * a.occ_cap];
            ev.target = tid2;
            double tp[D];
            load_pos<D>(a, tid2, tp);
            separation_vector<D>(a, sl.apos, tp, s);
            double q3[3] = {0, 0, 0};
            for (int d = 0; d < D; d++) q3[d] = s[(sl.adir + d) % D];
            const double c1 = T.use_charge ? charge_of(a, act) : 1.0;
            const double c2 = T.use_charge ? charge_of(a, tid2) : 1.0;
            const double der = isnan(ewald_pre) ? ewald_derivative_call(a.ewald, q3[0], q3[1], q3[2]) : ewald_pre;
            const double real = T.par[3] * c1 * c2 * der * sl.aspeed;
            if (real > 0) {
                double u = words_to_random(ring_word(a, st.cursor), ring_word(a, st.cursor + 1));
                st.cursor += 2;
                u = 0 + (sl.cv_rate - 0) * u;
                if (u < real) lift = true;
            }
        }
        store_pos<D>(a, act, sl.apos);
        st.at = sl.astamp;
        if (lift) {
            st.est[0] += s[sl.adir];
            st.est[4] += 1.0;
            st.active = ev.target;
            ev.flags |= 1;
        } else {
            st.est[6] += 1.0;
        }
        st.est[2] += 1.0;
        break;
    }
    case ECMC_TAG_CELL_BOUNDARY: {
#pragma unroll
        for (int d = 0; d < D; d++) p[d] = sl.apos[d];
        slice_pos<D>(a, p, sl.adir, sl.aspeed, sl.astamp, sl.t);
        p[sl.cb_dir] = sl.cb_boundary;
        store_pos<D>(a, act, p);
        st.at = sl.t;
        st.est[3] += 1.0;
        break;
    }
    case ECMC_TAG_SAMPLING:
    case ECMC_TAG_END_OF_RUN: {
        load_pos<D>(a, act, p);
        slice_pos<D>(a, p, st.adir, st.aspeed, st.at, sl.t);
        store_pos<D>(a, act, p);
        st.at = sl.t;
        if (T.kind == ECMC_TAG_END_OF_RUN) st.done = 1;
        break;
    }
    case ECMC_TAG_END_OF_CHAIN: {
        load_pos<D>(a, act, p);
        slice_pos<D>(a, p, st.adir, st.aspeed, st.at, sl.t);
        store_pos<D>(a, act, p);
        st.at = sl.t;
        st.eoc_last = sl.t;
        ev.target = sl.target;
        st.est[1] += T.value * st.aspeed;
        st.est[5] += 1.0;
        st.adir = (st.adir + 1) % D;
        if (sl.target != act) ev.flags |= 1;
        st.active = sl.target;
        break;
    }
    case ECMC_TAG_START_OF_RUN:
        st.active = T.ivalue2;
        st.adir = T.ivalue;
        st.aspeed = T.value;
        st.at.q = 0.0;
        st.at.r = 0.0;
        break;
    default:
        st.err = ECMC_ERR_CONFIGURATION;
    }
    ev.pos[0] = ev.pos[1] = ev.pos[2] = 0.0;
    if (act >= 0) {
        double q[D];
        load_pos<D>(a, act, q);
        for (int d = 0; d < D; d++) ev.pos[d] = q[d];
    }
}

// One workgroup runs one system's legs (the whole leg loop of the reference's mediator).
template <int D, int FM>
__device__ __forceinline__ void chain_body(const ChainArgs &a) {
    __shared__ ChainState st;
    __shared__ ecmc_tagger_desc tg[ECMC_MAX_TAGGERS];
    __shared__ uint32_t mt[624];
    __shared__ LegShared ls;
    int &flag = ls.flag;
    const int tid = threadIdx.x;
    // the positions and the MT ring move into LDS for the launch when the host sized it so (the
    // accessors load_pos, store_pos, ring_word and ring_ensure follow the flags)
    if (a.lds_pos) {
        double *lp = reinterpret_cast<double *>(chain_dyn_lds);
        for (int i = tid; i < D * a.n; i += NT) lp[i] = a.pos[i];
    }
    if (a.lds_ring) {
        uint32_t *lr = reinterpret_cast<uint32_t *>(chain_dyn_lds + a.lds_ring_off);
        for (int i = tid; i < a.ring_words; i += NT) lr[i] = a.ring[i];
    }
    if (a.lds_occn_off >= 0) {
        int32_t *lc = occn_base(a);
        for (int i = tid; i < a.ncells; i += NT) lc[i] = a.occ_n[i];
    }
    if (a.lds_charge_off >= 0) {
        double *lq = reinterpret_cast<double *>(chain_dyn_lds + a.lds_charge_off);
        for (int i = tid; i < a.n; i += NT) lq[i] = a.charge[i];
    }
    if (a.lds_occ_off >= 0) {
        int32_t *lo = occ_base(a);
        for (int i = tid; i < a.ncells * a.occ_cap; i += NT) lo[i] = a.occ[i];
    }
    if (tid == 0) st = *a.st;
    for (int i = tid; i < a.n_taggers; i += NT) tg[i] = a.tg[i];
    for (int i = tid; i < 624; i += NT) mt[i] = a.mt[i];
    __syncthreads();
    int64_t k = 0;
    while (k < a.max_events) {
        if (st.done || st.err) break;
        // every wave takes its branch decision before thread 0 changes the flags inside it
        const int started = st.started, pending = st.pending_creates;
        __syncthreads();
        if (!started) {
            if (tid == 0) {
                int sor = -1;
                for (int t = 0; t < a.n_taggers; t++)
                    if (tg[t].kind == ECMC_TAG_START_OF_RUN) sor = t;
                if (sor < 0) st.err = ECMC_ERR_CONFIGURATION;
                else create_single<D>(a, st, tg[sor], sor);
                st.started = 1;
                st.ncand = 0;
                st.chash = 0;
            }
            __syncthreads();
        } else if (pending) {
            if (a.occupancy) occupancy_update<D>(a, st, ls.scan);
            if (tid == 0) {
                st.ncand = 0;
                st.chash = 0;
            }
            __syncthreads();
            const ecmc_tagger_desc &w = tg[st.last_tagger];
            for (int ci = 0; ci < w.n_create && !st.err; ci++) {
                const int t = w.create[ci];
                create_tagger<D, FM>(a, st, mt, tg[t], t, ls);
            }
            if (tid == 0) st.pending_creates = 0;
            __syncthreads();
            if (st.err) break;
        }
        // earliest stored event (ties: first pushed): lane t of the first wavefront holds slot t,
        // a butterfly over the ECMC_MAX_TAGGERS lanes keeps the least (time, push counter) of the
        // valid slots (push counters are distinct, so the result is the sequential scan's)
        static_assert(ECMC_MAX_TAGGERS == 16, "the select butterfly spans 16 lanes");
        if (tid < 64) {
            const Slot &sl = st.slot[tid < ECMC_MAX_TAGGERS ? tid : 0];
            const bool v = tid < a.n_taggers && sl.valid;
            double q = v ? sl.t.q : 0.0, r = v ? sl.t.r : 0.0;
            int64_t sq = v ? sl.seq : 0;
            int w = v ? tid : -1;
#pragma unroll
            for (int off = 1; off < ECMC_MAX_TAGGERS; off <<= 1) {
                const double oq = __shfl_xor(q, off, 64), orr = __shfl_xor(r, off, 64);
                const int64_t os = __shfl_xor((long long)sq, off, 64);
                const int ow = __shfl_xor(w, off, 64);
                if (ow >= 0 && (w < 0 || oq < q || (oq == q && (orr < r || (orr == r && os < sq))))) {
                    q = oq;
                    r = orr;
                    sq = os;
                    w = ow;
                }
            }
            if (tid == 0) {
                if (w < 0) st.err = ECMC_ERR_SCHEDULER;
                else if (time_lt(st.slot[w].t, st.last)) st.err = ECMC_ERR_SCHEDULER;
                else st.last = st.slot[w].t;
                flag = w;
            }
        }
        __syncthreads();
        if (st.err) break;
        const int w = flag;
        if (commit_thins(tg[w])) ring_ensure(st, mt, a, 2);
        double ewald_pre = NAN;
        if ((FM & FEAT_EWALD) && tid < 64) {
            double q3[3];
            if (thinning_separation<D>(a, st, tg[w], w, q3)) ewald_pre = ewald_derivative_wave(a.ewald, a.ewald_codes, a.ewald_ncodes, q3[0], q3[1], q3[2]);
        }
        if (tid == 0) {
            ecmc_event ev;
            commit_event<D, FM>(a, st, tg[w], w, ev, ewald_pre);
            const ecmc_tagger_desc &T = tg[w];
            // trash by tag
            for (int i = 0; i < T.n_trash; i++) st.slot[T.trash[i]].valid = 0;
            if (a.log) a.log[k] = ev;
            st.pending_creates = 1;
            st.last_tagger = w;
            if (a.stop_mask & (1 << w)) st.stop_reason = 1;
        }
        k++;
        __syncthreads();
        if (st.stop_reason) break;
    }
    __syncthreads();
    if (tid == 0) {
        st.stop_reason = 0;
        *a.st = st;
        *a.n_out = k;
    }
    for (int i = tid; i < 624; i += NT) a.mt[i] = mt[i];
    if (a.lds_pos) {
        const double *lp = reinterpret_cast<const double *>(chain_dyn_lds);
        for (int i = tid; i < D * a.n; i += NT) a.pos[i] = lp[i];
    }
    if (a.lds_ring) {
        const uint32_t *lr = reinterpret_cast<const uint32_t *>(chain_dyn_lds + a.lds_ring_off);
        for (int i = tid; i < a.ring_words; i += NT) a.ring[i] = lr[i];
    }
    if (a.lds_occn_off >= 0) {
        const int32_t *lc = occn_base(a);
        for (int i = tid; i < a.ncells; i += NT) a.occ_n[i] = lc[i];
    }
    if (a.lds_occ_off >= 0) {
        const int32_t *lo = occ_base(a);
        for (int i = tid; i < a.ncells * a.occ_cap; i += NT) a.occ[i] = lo[i];
    }
}

// One system; its arguments in device memory (s->d_args), read at a uniform address.
template <int D, int FM>
__global__ __launch_bounds__(NT) void chain_kernel(const ChainArgs *__restrict__ args) {
    chain_body<D, FM>(*args);
}

// One step of a leg driven from the host, for the set-level event handler (the mediator and
// its scheduler stay on the host, the state and the candidate sets on the device):
//   op 0  the start of a leg: the cell occupancy follows the active unit and the candidate
//         statistics restart (the activator's internal-state update, tag_activator.py:205-226);
//   op 1  send_event_time of tagger t's handlers: its candidates' times, the earliest kept in
//         slot t (time, target, in-state copies);
//   op 2  send_out_state of tagger t's stored event, committed to the state; *out receives it.
template <int D, int FM>
__global__ __launch_bounds__(NT) void chain_step_kernel(const ChainArgs *__restrict__ args, int op, int t,
                                                        ecmc_event *out) {
    const ChainArgs &a = *args;
    __shared__ ChainState st;
    __shared__ ecmc_tagger_desc tg[ECMC_MAX_TAGGERS];
    __shared__ uint32_t mt[624];
    __shared__ LegShared ls;
    const int tid = threadIdx.x;
    if (tid == 0) st = *a.st;
    for (int i = tid; i < a.n_taggers; i += NT) tg[i] = a.tg[i];
    for (int i = tid; i < 624; i += NT) mt[i] = a.mt[i];
    __syncthreads();
    if (op == 0) {
        if (a.occupancy && st.active >= 0) occupancy_update<D>(a, st, ls.scan);
        if (tid == 0) {
            st.ncand = 0;
            st.chash = 0;
        }
        __syncthreads();
    } else if (op == 1) {
        // the occupancy follows the active unit here too (a no-op after op 0 in the same leg), so
        // that a cell tagger never reads the cell of a unit that stopped being active
        if (a.occupancy && st.active >= 0) occupancy_update<D>(a, st, ls.scan);
        create_tagger<D, FM>(a, st, mt, tg[t], t, ls);
    } else {
        if (commit_thins(tg[t])) ring_ensure(st, mt, a, 2);
        double ewald_pre = NAN;
        if ((FM & FEAT_EWALD) && tid < 64) {
            double q3[3];
            if (thinning_separation<D>(a, st, tg[t], t, q3)) ewald_pre = ewald_derivative_wave(a.ewald, a.ewald_codes, a.ewald_ncodes, q3[0], q3[1], q3[2]);
        }
        if (tid == 0) {
            ecmc_event ev;
            commit_event<D, FM>(a, st, tg[t], t, ev, ewald_pre);
            st.last = st.slot[t].t;
            *out = ev;
        }
    }
    __syncthreads();
    if (tid == 0) *a.st = st;
    for (int i = tid; i < 624; i += NT) a.mt[i] = mt[i];
}

// Independent systems in one launch: workgroup b runs system b (its own positions, cell
// occupancy, random stream and state), exactly as chain_kernel runs it alone.
template <int D, int FM>
__global__ __launch_bounds__(NT) void chain_many_kernel(const ChainArgs *__restrict__ args) {
    chain_body<D, FM>(args[blockIdx.x]);   // the arguments stay in global memory, read at a uniform address
}

// InnerPointEstimator.derivative_bound (inner_point_estimator.py:108-171) for the cell
// separations of cell_bounding_potential.py:100-126: one workgroup per (cell, direction)
// evaluates the merged-image Coulomb derivative on the (pps + 1)^3 grid spanning the corners
// (corners[6c .. 6c+6] = lower[3], upper[3]; NaN marks a nearby separation).
__global__ __launch_bounds__(NT) void cell_bound_kernel(Ewald ew, const double *corners, double L, double pot_pref,
                                                        double target_charge, int pps, double est_pref, double emp,
                                                        double *out) {
    __shared__ double smax[NT], smin[NT];
    const int c = blockIdx.x / 3, dir = blockIdx.x % 3, t = threadIdx.x;
    const double *cr = corners + (size_t)c * 6;
    double *o = out + (size_t)blockIdx.x * 2;
    if (isnan(cr[0])) {
        if (t == 0) { o[0] = NAN; o[1] = NAN; }
        return;
    }
    const int np = pps + 1, total = np * np * np;
    const double half = L / 2.0;
    double mx = -INFINITY, mn = INFINITY;
    for (int k = t; k < total; k += NT) {
        const int ix = k / (np * np), iy = (k / np) % np, iz = k % np;
        double sep[3];
        sep[0] = cr[0] + (cr[3] - cr[0]) * ix / pps;
        sep[1] = cr[1] + (cr[4] - cr[1]) * iy / pps;
        sep[2] = cr[2] + (cr[5] - cr[2]) * iz / pps;
#pragma unroll
        for (int d = 0; d < 3; d++) sep[d] = py_mod(sep[d] + half, L) - half;   // correct_separation
        double der = pot_pref * 1.0 * target_charge * ewald_derivative(ew, sep[dir], sep[(dir + 1) % 3], sep[(dir + 2) % 3]);
        der = der * 1.0;
        if (der > mx) mx = der;
        if (der < mn) mn = der;
    }
    smax[t] = mx;
    smin[t] = mn;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (t < w) {
            if (smax[t + w] > smax[t]) smax[t] = smax[t + w];
            if (smin[t + w] < smin[t]) smin[t] = smin[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        double upper = smax[0], lower = smin[0];
        upper = upper > 0.0 ? upper * est_pref : upper / est_pref;
        lower = lower > 0.0 ? lower / est_pref : lower * est_pref;
        o[0] = upper < emp ? upper : emp;
        o[1] = lower > -emp ? lower : -emp;
    }
}

// ---- host side --------------------------------------------------------------------------
static void mt_seed_host(uint64_t seed, uint32_t *mt) {
    uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
    int klen = key[1] ? 2 : 1;
    mt[0] = 19650218u;
    for (int i = 1; i < 624; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    int i = 1, j = 0;
    for (int k = 624 > klen ? 624 : klen; k; k--) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
        i++; j++;
        if (i >= 624) { mt[0] = mt[623]; i = 1; }
        if (j >= klen) j = 0;
    }
    for (int k = 623; k; k--) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        i++;
        if (i >= 624) { mt[0] = mt[623]; i = 1; }
    }
    mt[0] = 0x80000000u;
}

}  // namespace ecmc

using namespace ecmc;

// Walker's table (walker.py:73-101): small and large stacks split at the mean (sequential sum
// of the rates), one (small, alias large) entry per step with the large item's residual pushed
// back by comparison with the mean; the leftovers must lie within 1e-6 of the mean.
static bool walker_table(const std::vector<int32_t> &items, const std::vector<double> &rates, CvEntry *out,
                         double &total_out, double &mean_out) {
    const int m = (int)items.size();
    double total = 0.0;
    for (int i = 0; i < m; i++) total += rates[i];
    const double mean = total / m;
    std::vector<std::pair<int32_t, double>> small, large;
    for (int i = 0; i < m; i++) {
        if (!(rates[i] >= 0.0)) return false;
        if (rates[i] > mean) large.emplace_back(items[i], rates[i]);
        else small.emplace_back(items[i], rates[i]);
    }
    int k = 0;
    while (!small.empty() && !large.empty()) {
        const std::pair<int32_t, double> sm = small.back();
        small.pop_back();
        std::pair<int32_t, double> lg = large.back();
        large.pop_back();
        out[k++] = CvEntry{sm.second, sm.first, lg.first};
        lg.second -= mean - sm.second;
        if (lg.second < mean) small.push_back(lg);
        else large.push_back(lg);
    }
    for (std::vector<std::pair<int32_t, double>> *rest : {&small, &large}) {
        while (!rest->empty()) {
            const std::pair<int32_t, double> it = rest->back();
            rest->pop_back();
            const double q = it.second / mean;
            if (!(1 - 1e-6 < q && q < 1 + 1e-6)) return false;
            out[k++] = CvEntry{mean, it.first, -1};
        }
    }
    total_out = total;
    mean_out = mean;
    return k == m;
}

struct ecmc_system {
    ecmc_system_desc desc;
    int dim, n, ncells, nb_stride, occ_cap, ring_words;
    std::vector<double> cmin;
    double *d_pos = nullptr, *d_charge = nullptr, *d_cmin = nullptr, *d_ewald = nullptr;
    int32_t *d_ewald_codes = nullptr;
    int32_t *d_nearby = nullptr, *d_nearby_n = nullptr, *d_occ = nullptr, *d_occ_n = nullptr;
    int32_t *d_sur_cell = nullptr, *d_sur_id = nullptr, *d_cand = nullptr;
    double *d_cbt = nullptr, *d_corners = nullptr;
    CvEntry *d_cvw = nullptr;
    uint32_t *d_mt = nullptr, *d_ring = nullptr;
    ChainState *d_st = nullptr;
    ecmc_tagger_desc *d_tg = nullptr;
    ecmc_event *d_log = nullptr;
    int64_t log_cap = 0;
    int64_t *d_nout = nullptr;
    ecmc_event *d_step_ev = nullptr;   // the event of the last host-driven commit
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int32_t stop_mask = 0;
    int device = 0;          // the device current at creation, which holds its memory
    size_t lds_bytes = 0;    // dynamic LDS of a launch (positions and ring resident there)
    int features = FEAT_ALL; // the kernel features the system uses (FEAT_*, potential bits)
    ChainArgs args;
    ChainArgs *d_args = nullptr;   // the arguments of the system's own launches, in device memory
};

// The kernel features a system's taggers use.
static int system_features(const ecmc_system_desc &d) {
    int f = 0;
    for (int t = 0; t < d.n_taggers; t++) {
        const ecmc_tagger_desc &tg = d.taggers[t];
        const bool pair = tg.kind == ECMC_TAG_PAIR_CELLS || tg.kind == ECMC_TAG_PAIR_ALL || tg.kind == ECMC_TAG_SURPLUS ||
                          tg.kind == ECMC_TAG_CELL_BOUNDING || tg.kind == ECMC_TAG_FACTOR_MAP;
        if (pair) {
            if (tg.event_handler == ECMC_EH_TWO_LEAF && tg.potential >= 0 && tg.potential < 9) f |= 1 << tg.potential;
            else if (tg.event_handler == ECMC_EH_TWO_LEAF_BOUNDING) f |= FEAT_COULOMB;
            else if (tg.event_handler == ECMC_EH_CELL_BOUNDING) f |= FEAT_CELL_BOUNDING;
            else f |= FEAT_ALL;
            if (tg.kind == ECMC_TAG_FACTOR_MAP) f |= FEAT_FMAP;
            if (tg.kind == ECMC_TAG_CELL_BOUNDING) f |= FEAT_CELL_BOUNDING;
        }
        if (tg.kind == ECMC_TAG_CELL_VETO) f |= FEAT_CELL_VETO;
    }
    return f;
}

// The kernel variants of the leg loop, in the order the host tries them (three dimensions; two
// dimensions run the full kernel): the first whose features cover a launch's runs it.
struct ChainVariant {
    int features;
    const void *one, *many;   // chain_kernel, chain_many_kernel
};
static const ChainVariant CHAIN_VARIANTS_3D[] = {
    {VARIANT_LJ, reinterpret_cast<const void *>(&chain_kernel<3, VARIANT_LJ>),
     reinterpret_cast<const void *>(&chain_many_kernel<3, VARIANT_LJ>)},
    {VARIANT_CHARGED_CB, reinterpret_cast<const void *>(&chain_kernel<3, VARIANT_CHARGED_CB>),
     reinterpret_cast<const void *>(&chain_many_kernel<3, VARIANT_CHARGED_CB>)},
    {VARIANT_CHARGED_CV, reinterpret_cast<const void *>(&chain_kernel<3, VARIANT_CHARGED_CV>),
     reinterpret_cast<const void *>(&chain_many_kernel<3, VARIANT_CHARGED_CV>)},
    {VARIANT_DIPOLES, reinterpret_cast<const void *>(&chain_kernel<3, VARIANT_DIPOLES>),
     reinterpret_cast<const void *>(&chain_many_kernel<3, VARIANT_DIPOLES>)},
    {FEAT_ALL, reinterpret_cast<const void *>(&chain_kernel<3, FEAT_ALL>),
     reinterpret_cast<const void *>(&chain_many_kernel<3, FEAT_ALL>)},
};
static const ChainVariant CHAIN_VARIANTS_2D[] = {
    {FEAT_ALL, reinterpret_cast<const void *>(&chain_kernel<2, FEAT_ALL>),
     reinterpret_cast<const void *>(&chain_many_kernel<2, FEAT_ALL>)},
};
static const ChainVariant &chain_variant(int dim, int features) {
    if (dim == 2) return CHAIN_VARIANTS_2D[0];
    for (const ChainVariant &v : CHAIN_VARIANTS_3D)
        if ((features & ~v.features) == 0) return v;
    return CHAIN_VARIANTS_3D[sizeof(CHAIN_VARIANTS_3D) / sizeof(CHAIN_VARIANTS_3D[0]) - 1];
}

// Dynamic LDS of the leg loop for a system of n units in dim dimensions with a ring of
// ring_words words: the positions first, then the ring, each only while the whole fits next to
// the kernel's static LDS within the device's per-workgroup limit.
static void plan_lds(ecmc_system *s) {
    int dev_max = 0;
    hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device);
    hipFuncAttributes fa;
    size_t stat = 16384;
    const hipError_t e = hipFuncGetAttributes(&fa, chain_variant(s->dim, FEAT_ALL).one);
    if (e == hipSuccess) stat = fa.sharedSizeBytes;
    const size_t budget = dev_max > (int)stat ? (size_t)dev_max - stat : 0;
    const size_t pos_b = ((size_t)s->dim * s->n * sizeof(double) + 15) & ~(size_t)15;
    const size_t ring_b = (size_t)s->ring_words * sizeof(uint32_t);
    ChainArgs &a = s->args;
    a.lds_pos = a.lds_ring = 0;
    a.lds_ring_off = 0;
    size_t used = 0;
    if (pos_b <= budget) {
        a.lds_pos = 1;
        used = pos_b;
    }
    if (used + ring_b <= budget) {
        a.lds_ring = 1;
        a.lds_ring_off = (int32_t)used;
        used += ring_b;
    }
    // the charges (read per candidate by the Coulomb taggers), the far-cell list of a
    // cell-bounding tagger (n entries at most), the occupant counts and the occupant lists, each
    // when it fits after the others
    a.lds_cand_off = a.lds_occn_off = a.lds_occ_off = a.lds_charge_off = -1;
    bool charged = false;
    for (int t = 0; t < s->desc.n_taggers; t++) charged = charged || s->desc.taggers[t].use_charge;
    const size_t charge_b = charged ? ((size_t)s->n * sizeof(double) + 15) & ~(size_t)15 : 0;
    if (charge_b && used + charge_b <= budget) {
        a.lds_charge_off = (int32_t)used;
        used += charge_b;
    }
    const size_t cand_b = s->d_cand ? ((size_t)s->n * sizeof(int32_t) + 15) & ~(size_t)15 : 0;
    if (cand_b && used + cand_b <= budget) {
        a.lds_cand_off = (int32_t)used;
        used += cand_b;
    }
    const size_t occn_b = ((size_t)s->ncells * sizeof(int32_t) + 15) & ~(size_t)15;
    if (used + occn_b <= budget) {
        a.lds_occn_off = (int32_t)used;
        used += occn_b;
    }
    const size_t occ_b = ((size_t)s->ncells * s->occ_cap * sizeof(int32_t) + 15) & ~(size_t)15;
    if (used + occ_b <= budget) {
        a.lds_occ_off = (int32_t)used;
        used += occ_b;
    }
    s->lds_bytes = used;
    if (used > 65536) {
        // every variant may run the system (a launch of several systems takes the variant that
        // covers all of them)
        bool ok = true;
        const ChainVariant *vs = s->dim == 2 ? CHAIN_VARIANTS_2D : CHAIN_VARIANTS_3D;
        const size_t nv = s->dim == 2 ? sizeof(CHAIN_VARIANTS_2D) / sizeof(ChainVariant) : sizeof(CHAIN_VARIANTS_3D) / sizeof(ChainVariant);
        for (size_t i = 0; i < nv; i++) {
            ok = ok && hipFuncSetAttribute(vs[i].one, hipFuncAttributeMaxDynamicSharedMemorySize, (int)used) == hipSuccess;
            ok = ok && hipFuncSetAttribute(vs[i].many, hipFuncAttributeMaxDynamicSharedMemorySize, (int)used) == hipSuccess;
        }
        if (!ok) {
            // keep within the default limit: the ring alone, or nothing
            (void)hipGetLastError();
            a.lds_pos = 0;
            a.lds_ring = ring_b <= 65536 && ring_b <= budget ? 1 : 0;
            a.lds_ring_off = 0;
            a.lds_cand_off = a.lds_occn_off = a.lds_occ_off = a.lds_charge_off = -1;
            s->lds_bytes = a.lds_ring ? ring_b : 0;
        }
    }
}

extern "C" int ecmc_system_set_stop_mask(ecmc_system *sys, int32_t mask) {
    if (!sys) return ecmc_fail(ECMC_ERR_ARGUMENT, "null system");
    sys->stop_mask = mask;
    return ECMC_OK;
}

// Cell-bound tables already built in this process, per device, keyed by everything they
// depend on (cell corners, box, potential and estimator parameters); the last eight are kept.
struct CellBoundCacheEntry {
    int device;
    std::vector<double> key, table;
};
static std::mutex g_cb_mutex;
static std::vector<CellBoundCacheEntry> g_cb_cache;

static bool cell_bound_cache_find(const std::vector<double> &key, std::vector<double> &table) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    std::lock_guard<std::mutex> lock(g_cb_mutex);
    for (const CellBoundCacheEntry &e : g_cb_cache)
        if (e.device == dev && e.key.size() == key.size() &&
            memcmp(e.key.data(), key.data(), sizeof(double) * key.size()) == 0) {
            table = e.table;
            return true;
        }
    return false;
}

static void cell_bound_cache_store(const std::vector<double> &key, const std::vector<double> &table) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lock(g_cb_mutex);
    if (g_cb_cache.size() >= 8) g_cb_cache.erase(g_cb_cache.begin());
    g_cb_cache.push_back({dev, key, table});
}

extern "C" int ecmc_system_create(const ecmc_system_desc *desc, const double *positions, const double *charges,
                                  ecmc_system **out) {
    if (!desc || !positions || !out) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    const int dim = desc->dim, n = desc->n;
    if (dim != 2 && dim != 3) return ecmc_fail(ECMC_ERR_CONFIGURATION, "dimension must be 2 or 3");
    if (desc->rng_skip < 0) return ecmc_fail(ECMC_ERR_CONFIGURATION, "rng_skip must be >= 0");
    if (n < 2) return ecmc_fail(ECMC_ERR_CONFIGURATION, "at least two units are required");
    if (desc->n_taggers < 1 || desc->n_taggers > ECMC_MAX_TAGGERS)
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "number of taggers out of range");
    if (desc->neighbor_layers < 0 || desc->neighbor_layers > 1)
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "neighbor_layers must be 0 or 1");
    const int P = desc->leaves_per_root > 1 ? desc->leaves_per_root : 1;
    if (P > ECMC_MAX_LEAVES) return ecmc_fail(ECMC_ERR_CONFIGURATION, "more leaf units per composite object than ECMC_MAX_LEAVES");
    if (n % P != 0) return ecmc_fail(ECMC_ERR_CONFIGURATION, "the number of units is not a multiple of leaves_per_root");
    int ewald_tagger = -1, cb_tagger = -1;
    int64_t max_cand = 0;   // the most pair candidates one tagger creates in a leg
    for (int t = 0; t < desc->n_taggers; t++) {
        const ecmc_tagger_desc &tg = desc->taggers[t];
        if (tg.kind < 0 || tg.kind > ECMC_TAG_FACTOR_MAP) return ecmc_fail(ECMC_ERR_CONFIGURATION, "unknown tagger kind");
        // with composite objects the cells hold leaf units (cell_level 2): the excluded-cells and
        // cell-boundary taggers work on leaf units as on point masses; the others are not built
        if (P > 1 && (tg.kind == ECMC_TAG_PAIR_ALL || tg.kind == ECMC_TAG_SURPLUS || tg.kind == ECMC_TAG_CELL_BOUNDING ||
                      tg.kind == ECMC_TAG_CELL_VETO))
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "with composite objects the pair taggers are factor-type-map and excluded-cells taggers");
        if (tg.kind == ECMC_TAG_FACTOR_MAP) {
            int fmax = 0;
            for (int k = 0; k < P; k++) {
                if (tg.fmap_n[k] < 0 || tg.fmap_n[k] > ECMC_MAX_FACTOR_MAP)
                    return ecmc_fail(ECMC_ERR_CONFIGURATION, "factor type map entries out of range");
                fmax = tg.fmap_n[k] > fmax ? tg.fmap_n[k] : fmax;
                for (int e = 0; e < tg.fmap_n[k]; e++) {
                    const int tl = tg.fmap_target[k][e];
                    const bool ok = tg.fmap_local ? (tl >= 0 && tl < P && tl != k) : (tl >= P && tl < 2 * P);
                    if (!ok) return ecmc_fail(ECMC_ERR_CONFIGURATION, "factor type map leaf number out of range");
                }
            }
            const int64_t m = tg.fmap_local ? fmax : (int64_t)(n / P - 1) * fmax;
            max_cand = m > max_cand ? m : max_cand;
            if (!tg.fmap_local && n / P < 2) return ecmc_fail(ECMC_ERR_CONFIGURATION, "a non-local factor needs two composite objects");
            if (tg.n_handlers > 0 && tg.n_handlers < m)
                return ecmc_fail(ECMC_ERR_TAG_ACTIVATOR, "not-running event handler list would run empty");
        }
        if ((tg.kind == ECMC_TAG_CELL_VETO) != (tg.event_handler == ECMC_EH_CELL_VETO))
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "the cell-veto handler runs under the cell-veto tagger only");
        if (tg.kind == ECMC_TAG_SURPLUS && desc->max_occupants <= 0)
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "the surplus tagger needs a bounded cell occupancy");
        if ((tg.kind == ECMC_TAG_CELL_BOUNDING) != (tg.event_handler == ECMC_EH_CELL_BOUNDING))
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "the cell-bounding handler runs under the cell-bounding tagger only");
        if (tg.kind == ECMC_TAG_CELL_BOUNDING || tg.kind == ECMC_TAG_CELL_VETO) {
            if (tg.potential != ECMC_POT_MERGED_IMAGE_COULOMB || tg.bounding_potential != ECMC_POT_CELL_BOUNDING || dim != 3)
                return ecmc_fail(ECMC_ERR_CONFIGURATION, "cell bounding needs merged-image Coulomb with the cell bounding potential in 3D");
            if (desc->max_occupants != 1)
                return ecmc_fail(ECMC_ERR_CONFIGURATION, "cell bounding needs maximum_number_occupants = 1");
            if (!(tg.bpar[0] > 0.0) || tg.bpar[1] < 1.0 || tg.bpar[1] > 64.0 || (tg.use_charge && tg.bpar[2] == 0.0))
                return ecmc_fail(ECMC_ERR_CONFIGURATION, "estimator prefactor, points_per_side or target_charge out of range");
            if (cb_tagger >= 0) {
                const ecmc_tagger_desc &o = desc->taggers[cb_tagger];
                for (int q = 0; q < 4; q++)
                    if (o.bpar[q] != tg.bpar[q] || o.par[q] != tg.par[q])
                        return ecmc_fail(ECMC_ERR_CONFIGURATION, "one cell-bounding parameter set per system");
            }
            cb_tagger = t;
            if (ewald_tagger >= 0) {
                const ecmc_tagger_desc &o = desc->taggers[ewald_tagger];
                if (o.par[0] != tg.par[0] || o.par[1] != tg.par[1] || o.par[2] != tg.par[2])
                    return ecmc_fail(ECMC_ERR_CONFIGURATION, "one merged-image Coulomb parameter set per system");
            }
            ewald_tagger = t;
        }
        if (tg.n_create < 0 || tg.n_create > ECMC_MAX_TAGGERS || tg.n_trash < 0 || tg.n_trash > ECMC_MAX_TAGGERS)
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "create/trash list length out of range");
        if (tg.kind == ECMC_TAG_START_OF_RUN &&
            (tg.ivalue < 0 || tg.ivalue >= dim || tg.ivalue2 < 0 || tg.ivalue2 >= n || !(tg.value > 0.0)))
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "initial direction, active identifier or speed out of range");
        if (tg.kind == ECMC_TAG_END_OF_CHAIN && !(tg.value > 0.0))
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "chain_time must be > 0.0");
        if (tg.kind == ECMC_TAG_SAMPLING && !(tg.value > 0.0))
            return ecmc_fail(ECMC_ERR_CONFIGURATION, "sampling_interval must be > 0.0");
        for (int i = 0; i < tg.n_create; i++)
            if (tg.create[i] < 0 || tg.create[i] >= desc->n_taggers) return ecmc_fail(ECMC_ERR_CONFIGURATION, "create tag out of range");
        for (int i = 0; i < tg.n_trash; i++)
            if (tg.trash[i] < 0 || tg.trash[i] >= desc->n_taggers) return ecmc_fail(ECMC_ERR_CONFIGURATION, "trash tag out of range");
        if (tg.kind == ECMC_TAG_PAIR_CELLS || tg.kind == ECMC_TAG_PAIR_ALL || tg.kind == ECMC_TAG_SURPLUS ||
            tg.kind == ECMC_TAG_FACTOR_MAP) {
            if (tg.event_handler == ECMC_EH_TWO_LEAF_BOUNDING) {
                if (tg.potential != ECMC_POT_MERGED_IMAGE_COULOMB || tg.bounding_potential != ECMC_POT_INVERSE_POWER_COULOMB_BOUND || dim != 3)
                    return ecmc_fail(ECMC_ERR_CONFIGURATION, "bounding handler needs merged-image Coulomb with the inverse-power Coulomb bound in 3D");
                if (ewald_tagger >= 0) {
                    const ecmc_tagger_desc &o = desc->taggers[ewald_tagger];
                    if (o.par[0] != tg.par[0] || o.par[1] != tg.par[1] || o.par[2] != tg.par[2])
                        return ecmc_fail(ECMC_ERR_CONFIGURATION, "one merged-image Coulomb parameter set per system");
                }
                ewald_tagger = t;
            } else if (tg.potential == ECMC_POT_MERGED_IMAGE_COULOMB || tg.potential == ECMC_POT_INVERSE_POWER_COULOMB_BOUND) {
                return ecmc_fail(ECMC_ERR_CONFIGURATION, "the Coulomb potentials are not invertible here");
            }
            if (tg.n_handlers > 0 && tg.n_handlers < n - 1 && tg.kind == ECMC_TAG_PAIR_ALL)
                return ecmc_fail(ECMC_ERR_TAG_ACTIVATOR, "not-running event handler list would run empty");
        }
    }
    ecmc_system *s = new ecmc_system();
    s->desc = *desc;
    s->features = system_features(*desc);
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return ecmc_fail(ECMC_ERR_HIP, "no current HIP device");
    }
    s->dim = dim;
    s->n = n;
    CellGrid grid;
    std::string err;
    if (!build_cell_grid(dim, desc->system_length, desc->cells_per_side, desc->neighbor_layers, grid, err)) {
        delete s;
        return ecmc_fail(ECMC_ERR_CONFIGURATION, err.c_str());
    }
    s->ncells = grid.count;
    s->nb_stride = grid.nb_stride;
    // initial occupancy in global-state order (single_active_cell_occupancy.py:111-121)
    std::vector<std::vector<int32_t>> lists(grid.count);
    std::vector<int32_t> sur_cell, sur_id;     // the surplus dict, flattened in its yield order
    for (int i = 0; i < n; i++) {
        int c = grid.position_to_cell(positions + (size_t)i * dim);
        if (c < 0) { delete s; return ecmc_fail(ECMC_ERR_STATE, "initial position outside of the cell system"); }
        if (desc->max_occupants <= 0 || (int)lists[c].size() < desc->max_occupants) {
            lists[c].push_back(i);
        } else {
            size_t at = sur_cell.size();
            for (size_t k = sur_cell.size(); k-- > 0;)
                if (sur_cell[k] == c) { at = k + 1; break; }
            sur_cell.insert(sur_cell.begin() + at, c);
            sur_id.insert(sur_id.begin() + at, i);
        }
    }
    size_t maxocc = 0;
    for (auto &l : lists) maxocc = l.size() > maxocc ? l.size() : maxocc;
    int cap = (int)(maxocc * 4 + 32);
    if (cap > n) cap = n;
    s->occ_cap = cap;
    std::vector<int32_t> occ((size_t)grid.count * cap, -1), occ_n(grid.count);
    for (int c = 0; c < grid.count; c++) {
        occ_n[c] = (int32_t)lists[c].size();
        for (size_t i = 0; i < lists[c].size(); i++) occ[(size_t)c * cap + i] = lists[c][i];
    }
    std::vector<double> pos_soa((size_t)dim * n);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < dim; d++) pos_soa[(size_t)d * n + i] = positions[(size_t)i * dim + d];
    std::vector<double> ch(n, 1.0);
    if (charges) for (int i = 0; i < n; i++) ch[i] = charges[i];
    // the candidates of one create: all other units, the surplus (at most n), the occupants of
    // the nearby cells (a cell holds at most occ_cap: an overflow stops the run) or the occupied
    // far cells
    for (int t = 0; t < desc->n_taggers; t++) {
        const int kind = desc->taggers[t].kind;
        int64_t m = 0;
        if (kind == ECMC_TAG_PAIR_ALL || kind == ECMC_TAG_SURPLUS) m = n;
        else if (kind == ECMC_TAG_PAIR_CELLS) m = (int64_t)grid.nb_stride * cap;
        else if (kind == ECMC_TAG_CELL_BOUNDING) m = std::max<int64_t>((int64_t)grid.nb_stride * cap, grid.count);
        max_cand = m > max_cand ? m : max_cand;
    }
    // the ring holds every word a create may read plus a block in production: a power of two, so
    // that word indices wrap with a mask
    {
        const int64_t need = 624 * ((2 * (int64_t)max_cand + 256) / 624 + 3);
        int64_t words = 1024;
        while (words < need) words <<= 1;
        if (words > (1ll << 30)) { delete s; return ecmc_fail(ECMC_ERR_CONFIGURATION, "candidate sets too large"); }
        s->ring_words = (int)words;
    }
    uint32_t mt[624];
    mt_seed_host(desc->seed, mt);
    ChainState st;
    memset(&st, 0, sizeof(st));
    st.cursor = (uint64_t)desc->rng_skip;   // words drawn before the run (random input handler)
    st.active = -1;
    st.aspeed = 1.0;
    st.occ_id = -1;
    st.occ_cell = -1;
    st.last.q = -INFINITY;
    st.last.r = -INFINITY;
    st.nsur = (int32_t)sur_cell.size();
    for (int t = 0; t < desc->n_taggers; t++) {
        st.slot[t].valid = 0;
        const ecmc_tagger_desc &tg = desc->taggers[t];
        if (tg.kind == ECMC_TAG_SAMPLING && tg.ivalue) {
            Time z = time_from_float(-tg.value);
            st.samp_time = z;
        }
    }
    std::vector<double> ewald;
    std::vector<int32_t> ewald_codes;
    ChainArgs &a = s->args;
    memset(&a, 0, sizeof(a));
    if (ewald_tagger >= 0) {
        const ecmc_tagger_desc &tg = desc->taggers[ewald_tagger];
        int fc = (int)tg.par[1], pc = (int)tg.par[2];
        double alpha = tg.par[0], L = desc->system_length;
        build_ewald_table(fc, alpha, L, ewald);
        ewald_codes = ewald_term_codes(fc, pc);
        a.ewald.fc = fc;
        a.ewald.pc = pc;
        a.ewald.L = L;
        a.ewald.aol = alpha / L;
        a.ewald.aol2 = alpha * alpha / (L * L);
        a.ewald.two_aol_rpi = 2.0 * alpha / (L * sqrt(M_PI));
        a.ewald.tpol = 2.0 * M_PI / L;
    }
#define ECMC_ALLOC(ptr, bytes) \
    if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) { ecmc_system_destroy(s); return ecmc_fail(ECMC_ERR_HIP, "hipMalloc failed"); }
#define ECMC_UP(dst, src, bytes) \
    if (hipMemcpy((dst), (src), (bytes), hipMemcpyHostToDevice) != hipSuccess) { ecmc_system_destroy(s); return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed"); }
    ECMC_ALLOC(s->d_pos, sizeof(double) * dim * n);
    ECMC_ALLOC(s->d_args, sizeof(ChainArgs));
    ECMC_ALLOC(s->d_charge, sizeof(double) * n);
    ECMC_ALLOC(s->d_cmin, sizeof(double) * grid.cmin.size());
    ECMC_ALLOC(s->d_nearby, sizeof(int32_t) * grid.nearby.size());
    ECMC_ALLOC(s->d_nearby_n, sizeof(int32_t) * grid.nearby_n.size());
    ECMC_ALLOC(s->d_occ, sizeof(int32_t) * occ.size());
    ECMC_ALLOC(s->d_occ_n, sizeof(int32_t) * occ_n.size());
    ECMC_ALLOC(s->d_mt, sizeof(uint32_t) * 624);
    ECMC_ALLOC(s->d_ring, sizeof(uint32_t) * s->ring_words);
    ECMC_ALLOC(s->d_st, sizeof(ChainState));
    ECMC_ALLOC(s->d_tg, sizeof(ecmc_tagger_desc) * ECMC_MAX_TAGGERS);
    ECMC_ALLOC(s->d_nout, sizeof(int64_t));
    ECMC_ALLOC(s->d_step_ev, sizeof(ecmc_event));
    if (!ewald.empty()) { ECMC_ALLOC(s->d_ewald, sizeof(double) * ewald.size()); }
    if (!ewald_codes.empty()) { ECMC_ALLOC(s->d_ewald_codes, sizeof(int32_t) * ewald_codes.size()); }
    ECMC_ALLOC(s->d_sur_cell, sizeof(int32_t) * n);
    ECMC_ALLOC(s->d_sur_id, sizeof(int32_t) * n);
    ECMC_ALLOC(s->d_cand, sizeof(int32_t) * n);
    if (!sur_cell.empty()) {
        ECMC_UP(s->d_sur_cell, sur_cell.data(), sizeof(int32_t) * sur_cell.size());
        ECMC_UP(s->d_sur_id, sur_id.data(), sizeof(int32_t) * sur_id.size());
    }
    ECMC_UP(s->d_pos, pos_soa.data(), sizeof(double) * dim * n);
    ECMC_UP(s->d_charge, ch.data(), sizeof(double) * n);
    ECMC_UP(s->d_cmin, grid.cmin.data(), sizeof(double) * grid.cmin.size());
    ECMC_UP(s->d_nearby, grid.nearby.data(), sizeof(int32_t) * grid.nearby.size());
    ECMC_UP(s->d_nearby_n, grid.nearby_n.data(), sizeof(int32_t) * grid.nearby_n.size());
    ECMC_UP(s->d_occ, occ.data(), sizeof(int32_t) * occ.size());
    ECMC_UP(s->d_occ_n, occ_n.data(), sizeof(int32_t) * occ_n.size());
    ECMC_UP(s->d_mt, mt, sizeof(uint32_t) * 624);
    ECMC_UP(s->d_st, &st, sizeof(ChainState));
    ECMC_UP(s->d_tg, desc->taggers, sizeof(ecmc_tagger_desc) * ECMC_MAX_TAGGERS);
    if (!ewald.empty()) { ECMC_UP(s->d_ewald, ewald.data(), sizeof(double) * ewald.size()); }
    if (!ewald_codes.empty()) { ECMC_UP(s->d_ewald_codes, ewald_codes.data(), sizeof(int32_t) * ewald_codes.size()); }
#undef ECMC_ALLOC
#undef ECMC_UP
    if (hipStreamCreate(&s->stream) != hipSuccess || hipEventCreate(&s->ev0) != hipSuccess ||
        hipEventCreate(&s->ev1) != hipSuccess) {
        ecmc_system_destroy(s);
        return ecmc_fail(ECMC_ERR_HIP, "stream/event creation failed");
    }
    a.n = n;
    a.leaves = P;
    for (int t = 0; t < desc->n_taggers; t++) {
        const int k = desc->taggers[t].kind;
        if (k == ECMC_TAG_PAIR_CELLS || k == ECMC_TAG_SURPLUS || k == ECMC_TAG_CELL_BOUNDING || k == ECMC_TAG_CELL_VETO)
            a.occupancy = 1;
    }
    a.ncells = grid.count;
    a.nb_stride = grid.nb_stride;
    a.occ_cap = cap;
    a.ring_words = s->ring_words;
    a.n_taggers = desc->n_taggers;
    a.L = desc->system_length;
    a.half = desc->system_length / 2.0;
    a.beta = desc->beta;
    for (int d = 0; d < 3; d++) {
        a.cps[d] = d < dim ? grid.cps[d] : 1;
        a.cum[d] = d < dim ? (int32_t)grid.cum[d] : 0;
        a.side[d] = d < dim ? grid.side[d] : 1.0;
    }
    a.pos = s->d_pos;
    a.charge = s->d_charge;
    a.cmin = s->d_cmin;
    a.nearby = s->d_nearby;
    a.nearby_n = s->d_nearby_n;
    a.occ = s->d_occ;
    a.occ_n = s->d_occ_n;
    a.mt = s->d_mt;
    a.ring = s->d_ring;
    a.st = s->d_st;
    a.tg = s->d_tg;
    a.ewald.f = s->d_ewald;
    a.ewald_codes = s->d_ewald_codes;
    a.ewald_ncodes = (int32_t)ewald_codes.size();
    a.n_out = s->d_nout;
    a.max_occ = desc->max_occupants;
    a.sur_cap = n;
    a.layers = desc->neighbor_layers;
    a.sur_cell = s->d_sur_cell;
    a.sur_id = s->d_sur_id;
    a.cand = s->d_cand;
    if (cb_tagger >= 0) {
        // cell-bound tables on the device (cell_bounding_potential.py:100-126)
        const ecmc_tagger_desc &tg = desc->taggers[cb_tagger];
        std::vector<double> corners((size_t)grid.count * 6);
        const int32_t *near0 = grid.nearby.data();
        for (int c = 0; c < grid.count; c++) {
            bool is_near = false;
            for (int k = 0; k < grid.nearby_n[0]; k++) if (near0[k] == c) is_near = true;
            for (int d = 0; d < 3; d++) {
                corners[(size_t)c * 6 + d] = is_near ? NAN : grid.cmin[(size_t)c * dim + d] - grid.cmax[d];
                corners[(size_t)c * 6 + 3 + d] = is_near ? NAN : grid.cmax[(size_t)c * dim + d] - grid.cmin[d];
            }
        }
        if (hipMalloc((void **)&s->d_cbt, sizeof(double) * grid.count * 6) != hipSuccess) {
            ecmc_system_destroy(s);
            return ecmc_fail(ECMC_ERR_HIP, "cell-bound table allocation failed");
        }
        // the table depends only on the cell geometry, the box, the potential and the
        // estimator: replicas of one system take it from the cache instead of rebuilding it
        std::vector<double> key(corners);
        key.push_back(desc->system_length);
        for (int q = 0; q < 4; q++) key.push_back(tg.par[q]);
        for (int q = 0; q < 4; q++) key.push_back(tg.bpar[q]);
        std::vector<double> table;
        if (cell_bound_cache_find(key, table)) {
            if (hipMemcpy(s->d_cbt, table.data(), sizeof(double) * table.size(), hipMemcpyHostToDevice) != hipSuccess) {
                ecmc_system_destroy(s);
                return ecmc_fail(ECMC_ERR_HIP, "cell-bound table upload failed");
            }
        } else {
            if (hipMalloc((void **)&s->d_corners, sizeof(double) * corners.size()) != hipSuccess ||
                hipMemcpy(s->d_corners, corners.data(), sizeof(double) * corners.size(), hipMemcpyHostToDevice) != hipSuccess) {
                ecmc_system_destroy(s);
                return ecmc_fail(ECMC_ERR_HIP, "cell-bound table allocation failed");
            }
            hipLaunchKernelGGL(cell_bound_kernel, dim3(grid.count * 3), dim3(NT), 0, s->stream, a.ewald, s->d_corners,
                               desc->system_length, tg.par[3], tg.bpar[2], (int)tg.bpar[1], tg.bpar[0], tg.bpar[3],
                               s->d_cbt);
            table.resize((size_t)grid.count * 6);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess ||
                hipMemcpy(table.data(), s->d_cbt, sizeof(double) * table.size(), hipMemcpyDeviceToHost) != hipSuccess) {
                ecmc_system_destroy(s);
                return ecmc_fail(ECMC_ERR_HIP, "cell-bound table kernel failed");
            }
            cell_bound_cache_store(key, table);
        }
        a.cbt = s->d_cbt;
        bool veto = false;
        for (int t = 0; t < desc->n_taggers; t++) veto = veto || desc->taggers[t].kind == ECMC_TAG_CELL_VETO;
        if (veto) {
            // cell_veto_event_handler.py:147-159: per direction, the non-nearby cells in
            // yield_cells() order with rates max(upper, 0) and max(-lower, 0)
            std::vector<int32_t> items;
            for (int c = 0; c < grid.count; c++)
                if (!std::isnan(table[(size_t)c * 6])) items.push_back(c);
            const int m = (int)items.size();
            std::vector<CvEntry> all((size_t)6 * m);
            std::vector<double> rates(m);
            for (int wt = 0; wt < 6 && m > 0; wt++) {
                const int d = wt / 2, b = wt % 2;
                for (int i = 0; i < m; i++) {
                    const double v = table[(size_t)(items[i] * 3 + d) * 2 + b];
                    const double r = b ? -v : v;
                    rates[i] = 0.0 > r ? 0.0 : r;
                }
                if (!walker_table(items, rates, &all[(size_t)wt * m], a.cv_total[wt], a.cv_mean[wt])) {
                    ecmc_system_destroy(s);
                    return ecmc_fail(ECMC_ERR_CONFIGURATION, "cell veto: Walker table failed (negative rate or leftover check)");
                }
            }
            if (m == 0) {
                ecmc_system_destroy(s);
                return ecmc_fail(ECMC_ERR_CONFIGURATION, "cell veto: no cell lies outside the nearby cells");
            }
            if (hipMalloc((void **)&s->d_cvw, sizeof(CvEntry) * all.size()) != hipSuccess ||
                hipMemcpy(s->d_cvw, all.data(), sizeof(CvEntry) * all.size(), hipMemcpyHostToDevice) != hipSuccess) {
                ecmc_system_destroy(s);
                return ecmc_fail(ECMC_ERR_HIP, "cell veto table upload failed");
            }
            a.cvw = s->d_cvw;
            a.cv_m = m;
        }
    }
    plan_lds(s);
    *out = s;
    return ECMC_OK;
}

extern "C" void ecmc_system_destroy(ecmc_system *s) {
    if (!s) return;
    hipFree(s->d_pos); hipFree(s->d_charge); hipFree(s->d_cmin); hipFree(s->d_nearby); hipFree(s->d_nearby_n);
    hipFree(s->d_occ); hipFree(s->d_occ_n); hipFree(s->d_mt); hipFree(s->d_ring); hipFree(s->d_st);
    hipFree(s->d_sur_cell); hipFree(s->d_sur_id); hipFree(s->d_cand); hipFree(s->d_cbt); hipFree(s->d_corners);
    hipFree(s->d_cvw);
    hipFree(s->d_tg); hipFree(s->d_nout); hipFree(s->d_log); hipFree(s->d_ewald); hipFree(s->d_ewald_codes); hipFree(s->d_step_ev); hipFree(s->d_args);
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

static int status_of_state(const ChainState &st) {
    switch (st.err) {
    case 0: return ECMC_OK;
    case ECMC_ERR_SCHEDULER: return ecmc_fail(ECMC_ERR_SCHEDULER, "SchedulerError: no event or decreasing event time");
    case ECMC_ERR_STATE: return ecmc_fail(ECMC_ERR_STATE, "a position left the cell system or an occupancy list overflowed");
    case ECMC_ERR_NUMERIC: return ecmc_fail(ECMC_ERR_NUMERIC, "ValueError: math domain error in a displacement");
    case ECMC_ERR_CONFIGURATION: return ecmc_fail(ECMC_ERR_CONFIGURATION, "end-of-chain time assertion or missing start of run");
    default: return ecmc_fail(st.err, "device error");
    }
}

extern "C" int ecmc_system_run(ecmc_system *s, int64_t max_events, ecmc_event *host_log, int64_t *n_committed) {
    if (!s || max_events < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (n_committed) *n_committed = 0;
    if (max_events == 0) return ECMC_OK;
    if (host_log && s->log_cap < max_events) {
        hipFree(s->d_log);
        s->d_log = nullptr;
        if (hipMalloc((void **)&s->d_log, sizeof(ecmc_event) * max_events) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "hipMalloc of the event log failed");
        s->log_cap = max_events;
    }
    ChainArgs a = s->args;
    a.log = host_log ? s->d_log : nullptr;
    a.max_events = max_events;
    a.stop_mask = s->stop_mask;
    hipEventRecord(s->ev0, s->stream);
    if (hipMemcpyAsync(s->d_args, &a, sizeof(ChainArgs), hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy of the launch arguments failed");
    {
        void *kargs[] = {&s->d_args};
        hipLaunchKernel(chain_variant(s->dim, s->features).one, dim3(1), dim3(NT), kargs, s->lds_bytes, s->stream);
    }
    hipEventRecord(s->ev1, s->stream);
    if (hipGetLastError() != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "kernel launch failed");
    if (hipStreamSynchronize(s->stream) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "kernel execution failed");
    float ms = 0.f;
    hipEventElapsedTime(&ms, s->ev0, s->ev1);
    s->last_ms = ms;
    int64_t k = 0;
    ChainState st;
    if (hipMemcpy(&k, s->d_nout, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&st, s->d_st, sizeof(ChainState), hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    if (host_log && k > 0 &&
        hipMemcpy(host_log, s->d_log, sizeof(ecmc_event) * k, hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy of the event log failed");
    if (n_committed) *n_committed = k;
    return status_of_state(st);
}

// The launch of independent systems with their arguments staged in `host_args` (count
// entries, filled here) and copied to the device array `d_args`: one workgroup per system on the
// first system's stream, then every system's committed legs and state are read back.
static int systems_run_staged(ecmc_system *const *systems, int32_t count, int64_t max_events, int64_t *n_committed,
                              ChainArgs *host_args, ChainArgs *d_args) {
    const int dim = systems[0]->dim;
    size_t lds = 0;
    for (int i = 0; i < count; i++) {
        ecmc_system *s = systems[i];
        if (s->lds_bytes > lds) lds = s->lds_bytes;
        host_args[i] = s->args;
        host_args[i].log = nullptr;
        host_args[i].max_events = max_events;
        host_args[i].stop_mask = s->stop_mask;
        if (n_committed) n_committed[i] = 0;
    }
    if (max_events == 0) return ECMC_OK;
    ecmc_system *s0 = systems[0];
    if (hipMemcpyAsync(d_args, host_args, sizeof(ChainArgs) * count, hipMemcpyHostToDevice, s0->stream) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy of the launch arguments failed");
    hipEventRecord(s0->ev0, s0->stream);
    int features = 0;
    for (int i = 0; i < count; i++) features |= systems[i]->features;
    void *kargs[] = {&d_args};
    hipLaunchKernel(chain_variant(dim, features).many, dim3(count), dim3(NT), kargs, lds, s0->stream);
    hipEventRecord(s0->ev1, s0->stream);
    const bool launched = hipGetLastError() == hipSuccess;
    const bool ran = hipStreamSynchronize(s0->stream) == hipSuccess;
    if (!launched) return ecmc_fail(ECMC_ERR_HIP, "kernel launch failed");
    if (!ran) return ecmc_fail(ECMC_ERR_HIP, "kernel execution failed");
    float ms = 0.f;
    hipEventElapsedTime(&ms, s0->ev0, s0->ev1);
    int rc = ECMC_OK;
    for (int i = 0; i < count; i++) {
        ecmc_system *s = systems[i];
        s->last_ms = ms;
        int64_t k = 0;
        ChainState st;
        if (hipMemcpy(&k, s->d_nout, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&st, s->d_st, sizeof(ChainState), hipMemcpyDeviceToHost) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
        if (n_committed) n_committed[i] = k;
        if (rc == ECMC_OK) rc = status_of_state(st);
    }
    return rc;
}

static int check_systems(ecmc_system *const *systems, int32_t count, int64_t max_events) {
    if (!systems || count < 1 || max_events < 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (!systems[0]) return ecmc_fail(ECMC_ERR_ARGUMENT, "null system");
    const int dim = systems[0]->dim;
    for (int i = 0; i < count; i++) {
        ecmc_system *s = systems[i];
        if (!s) return ecmc_fail(ECMC_ERR_ARGUMENT, "null system");
        if (s->dim != dim) return ecmc_fail(ECMC_ERR_ARGUMENT, "all systems of one launch must have one dimension");
        if (s->device != systems[0]->device)
            return ecmc_fail(ECMC_ERR_ARGUMENT, "all systems of one launch must live on one device");
        for (int k = 0; k < i; k++)
            if (systems[k] == s) return ecmc_fail(ECMC_ERR_ARGUMENT, "a system appears twice in one launch");
    }
    return ECMC_OK;
}

extern "C" int ecmc_systems_run(ecmc_system *const *systems, int32_t count, int64_t max_events,
                                int64_t *n_committed) {
    int rc = check_systems(systems, count, max_events);
    if (rc) return rc;
    std::vector<ChainArgs> args((size_t)count);
    ChainArgs *d_args = nullptr;
    if (hipMalloc((void **)&d_args, sizeof(ChainArgs) * count) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMalloc of the launch arguments failed");
    rc = systems_run_staged(systems, count, max_events, n_committed, args.data(), d_args);
    hipFree(d_args);
    return rc;
}

// ---- independent systems without allocation -------------------------------------------------
// The launch arguments of ecmc_systems_run_noalloc live in a fixed device array of the library
// (each GPU the library runs on holds its own copy until the process ends) and are staged in a
// fixed host array; one process-wide lock serialises the calls, so calls from one process never
// share the arrays at the same time.  The array's device address comes from
// hipGetSymbolAddress once per device (the name itself is not a device address in host code).
static_assert(sizeof(ChainArgs) <= 512, "the no-allocation argument array is sized for 512 B per system");
static __device__ __align__(256) ChainArgs scratch[ECMC_SYSTEMS_NOALLOC_MAX];
static ChainArgs g_noalloc_host[ECMC_SYSTEMS_NOALLOC_MAX];
constexpr int NOALLOC_DEVICES = 64;
static void *g_scratch_ptr[NOALLOC_DEVICES];
static std::mutex g_noalloc_mutex;

extern "C" int ecmc_systems_run_noalloc(ecmc_system *const *systems, int32_t count, int64_t max_events,
                                        int64_t *n_committed) {
    if (count > ECMC_SYSTEMS_NOALLOC_MAX)
        return ecmc_fail(ECMC_ERR_ARGUMENT, "more systems than ECMC_SYSTEMS_NOALLOC_MAX in one no-allocation launch");
    int rc = check_systems(systems, count, max_events);
    if (rc) return rc;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "no current HIP device");
    if (dev != systems[0]->device)
        return ecmc_fail(ECMC_ERR_ARGUMENT, "the systems' device must be the current device");
    if (dev < 0 || dev >= NOALLOC_DEVICES) return ecmc_fail(ECMC_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_scratch_ptr[dev]) {
        // set-up on the first call on this device, with the device current
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return ecmc_fail(ECMC_ERR_HIP, "hipGetSymbolAddress of the argument array failed");
        g_scratch_ptr[dev] = p;
    }
    return systems_run_staged(systems, count, max_events, n_committed, g_noalloc_host,
                              (ChainArgs *)g_scratch_ptr[dev]);
}

// ---- the leg driven step by step from the host (set-level event handler) -------------------
static int system_step(ecmc_system *s, int op, int tagger, ChainState &st) {
    if (!s) return ecmc_fail(ECMC_ERR_ARGUMENT, "null system");
    if (op != 0 && (tagger < 0 || tagger >= s->desc.n_taggers)) return ecmc_fail(ECMC_ERR_ARGUMENT, "no such tagger");
    ChainArgs sa = s->args;   // one step works on global memory
    sa.lds_pos = sa.lds_ring = 0;
    sa.lds_cand_off = sa.lds_occn_off = sa.lds_occ_off = sa.lds_charge_off = -1;
    if (hipMemcpyAsync(s->d_args, &sa, sizeof(ChainArgs), hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy of the launch arguments failed");
    if (s->dim == 2) hipLaunchKernelGGL((chain_step_kernel<2, FEAT_ALL>), dim3(1), dim3(NT), 0, s->stream, s->d_args, op, tagger, s->d_step_ev);
    else hipLaunchKernelGGL((chain_step_kernel<3, FEAT_ALL>), dim3(1), dim3(NT), 0, s->stream, s->d_args, op, tagger, s->d_step_ev);
    if (hipGetLastError() != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "kernel launch failed");
    if (hipStreamSynchronize(s->stream) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "kernel execution failed");
    if (hipMemcpy(&st, s->d_st, sizeof(ChainState), hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    return status_of_state(st);
}

extern "C" int ecmc_system_begin_leg(ecmc_system *s) {
    ChainState st;
    return system_step(s, 0, 0, st);
}

extern "C" int ecmc_system_send_event_time(ecmc_system *s, int32_t tagger, double *tq, double *tr, int32_t *target,
                                           int32_t *valid) {
    ChainState st;
    const int rc = system_step(s, 1, tagger, st);
    if (rc) return rc;
    const Slot &sl = st.slot[tagger];
    if (tq) *tq = sl.valid ? sl.t.q : INFINITY;
    if (tr) *tr = sl.valid ? sl.t.r : INFINITY;
    if (target) *target = sl.valid ? sl.target : -1;
    if (valid) *valid = sl.valid;
    return ECMC_OK;
}

extern "C" int ecmc_system_send_out_state(ecmc_system *s, int32_t tagger, ecmc_event *event) {
    ChainState st;
    if (s && tagger >= 0 && tagger < s->desc.n_taggers) {
        // the handler must hold an event (the host's scheduler never returns one without)
        Slot sl;
        if (hipMemcpy(&sl, (const char *)s->d_st + offsetof(ChainState, slot) + sizeof(Slot) * tagger, sizeof(Slot),
                      hipMemcpyDeviceToHost) != hipSuccess)
            return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
        if (!sl.valid) return ecmc_fail(ECMC_ERR_SCHEDULER, "the event handler holds no event");
    }
    const int rc = system_step(s, 2, tagger, st);
    if (rc) return rc;
    if (event && hipMemcpy(event, s->d_step_ev, sizeof(ecmc_event), hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    return ECMC_OK;
}

extern "C" int ecmc_system_positions(const ecmc_system *s, double *out) {
    if (!s || !out) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    std::vector<double> soa((size_t)s->dim * s->n);
    if (hipMemcpy(soa.data(), s->d_pos, sizeof(double) * soa.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    for (int i = 0; i < s->n; i++)
        for (int d = 0; d < s->dim; d++) out[(size_t)i * s->dim + d] = soa[(size_t)d * s->n + i];
    return ECMC_OK;
}

static int read_state(const ecmc_system *s, ChainState &st) {
    if (hipMemcpy(&st, s->d_st, sizeof(ChainState), hipMemcpyDeviceToHost) != hipSuccess)
        return ecmc_fail(ECMC_ERR_HIP, "hipMemcpy failed");
    return ECMC_OK;
}

extern "C" int ecmc_system_active(const ecmc_system *s, int32_t *active, int32_t *direction, double *tq, double *tr,
                                  int32_t *done) {
    if (!s) return ecmc_fail(ECMC_ERR_ARGUMENT, "null system");
    ChainState st;
    int rc = read_state(s, st);
    if (rc) return rc;
    if (active) *active = st.active;
    if (direction) *direction = st.adir;
    if (tq) *tq = st.at.q;
    if (tr) *tr = st.at.r;
    if (done) *done = st.done;
    return ECMC_OK;
}

extern "C" int ecmc_system_estimators(const ecmc_system *s, double *est8) {
    if (!s || !est8) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    ChainState st;
    int rc = read_state(s, st);
    if (rc) return rc;
    memcpy(est8, st.est, sizeof(st.est));
    return ECMC_OK;
}

extern "C" int ecmc_system_rng_words(const ecmc_system *s, int64_t *words) {
    if (!s || !words) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    ChainState st;
    int rc = read_state(s, st);
    if (rc) return rc;
    *words = (int64_t)st.cursor;
    return ECMC_OK;
}

extern "C" int ecmc_system_launch_lds(const ecmc_system *s, int64_t *bytes, int32_t *positions, int32_t *ring) {
    if (!s) return ecmc_fail(ECMC_ERR_ARGUMENT, "null system");
    if (bytes) *bytes = (int64_t)s->lds_bytes;
    if (positions) *positions = s->args.lds_pos;
    if (ring) *ring = s->args.lds_ring;
    return ECMC_OK;
}

extern "C" int ecmc_system_kernel_variant(const ecmc_system *s, int32_t *features, int32_t *variant) {
    if (!s) return ecmc_fail(ECMC_ERR_ARGUMENT, "null system");
    if (features) *features = s->features;
    if (variant) *variant = chain_variant(s->dim, s->features).features;
    return ECMC_OK;
}

extern "C" int ecmc_desc_kernel_variant(const ecmc_system_desc *desc, int32_t *features, int32_t *variant) {
    if (!desc) return ecmc_fail(ECMC_ERR_ARGUMENT, "null descriptor");
    if (desc->dim != 2 && desc->dim != 3) return ecmc_fail(ECMC_ERR_CONFIGURATION, "dimension must be 2 or 3");
    if (desc->n_taggers < 0 || desc->n_taggers > ECMC_MAX_TAGGERS)
        return ecmc_fail(ECMC_ERR_CONFIGURATION, "number of taggers out of range");
    const int f = system_features(*desc);
    if (features) *features = f;
    if (variant) *variant = chain_variant(desc->dim, f).features;
    return ECMC_OK;
}

extern "C" int ecmc_system_last_kernel_ms(const ecmc_system *s, double *ms) {
    if (!s || !ms) return ecmc_fail(ECMC_ERR_ARGUMENT, "null argument");
    *ms = s->last_ms;
    return ECMC_OK;
}

// ---- MT19937 words on the device (testing the parallel twist against CPython) ----------
__global__ __launch_bounds__(NT) void mt_words_kernel(uint32_t *state, uint32_t *out, int64_t n) {
    __shared__ uint32_t mt[624];
    for (int i = threadIdx.x; i < 624; i += NT) mt[i] = state[i];
    __syncthreads();
    for (int64_t produced = 0; produced < n; produced += 624) {
        mt_generate_block(mt, out + produced, ~0ull, 0);
    }
}

extern "C" int ecmc_mt_words(uint64_t seed, int64_t n, uint32_t *d_out, void *stream) {
    if (n < 0 || (n > 0 && !d_out)) return ecmc_fail(ECMC_ERR_ARGUMENT, "bad argument");
    if (n % 624 != 0) return ecmc_fail(ECMC_ERR_ARGUMENT, "n must be a multiple of 624");
    if (n == 0) return ECMC_OK;
    uint32_t mt[624];
    mt_seed_host(seed, mt);
    uint32_t *d_state = nullptr;
    if (hipMalloc((void **)&d_state, sizeof(mt)) != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "hipMalloc failed");
    hipStream_t st = (hipStream_t)stream;
    hipMemcpyAsync(d_state, mt, sizeof(mt), hipMemcpyHostToDevice, st);
    hipLaunchKernelGGL(mt_words_kernel, dim3(1), dim3(NT), 0, st, d_state, d_out, n);
    hipError_t e = hipStreamSynchronize(st);
    hipFree(d_state);
    if (e != hipSuccess) return ecmc_fail(ECMC_ERR_HIP, "mt kernel failed");
    return ECMC_OK;
}
