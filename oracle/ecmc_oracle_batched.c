/*
 * ecmc_oracle_batched.c -- CPU restatement of the batched-replica ECMC of the product
 * (TEST INFRASTRUCTURE ONLY, see ecmc_oracle.h).
 *
 * The batched form has no reference counterpart as a whole; it is the reference's chain
 * (straight chains of fixed length cycling +x,+y,+z, pair factors inverted exactly,
 * cell-boundary events of CuboidPeriodicCells, a random new active unit at the end of each
 * chain) with these defined differences, which the device kernel shares:
 *   - the potential is the truncated-and-shifted inverse power (two factors for LJ 12-6);
 *   - random numbers are Philox4x32-10 (Salmon et al., SC'11) keyed by (replica, seed)
 *     with counter (leg, unit) -- so they do not depend on the order candidates are visited;
 *   - time is the displacement along the chain (speed 1) in one double per chain; the
 *     minimum image is s = t - a folded once into [-L/2, L/2);
 *   - equal event displacements are resolved by (pair < cell boundary < end of chain), then
 *     target identifier, then factor index.
 * Every candidate is evaluated here in full double precision with correctly rounded pow
 * and log (quad precision), which is what the device's FP32-filtered evaluation must
 * reproduce bit for bit.
 */
#include "ecmc_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- Philox4x32-10 ------------------------------------------------------------------ */
void ora_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double words_to_unit(uint32_t a, uint32_t b) {
    return ((a >> 5) * 67108864.0 + (b >> 6)) * (1.0 / 9007199254740992.0);
}

/* ---- cell grid with float-exact lower bounds (as CuboidCells) -------------------------- */
typedef struct {
    int cps;         /* cells per side, 0 = all pairs */
    double side;
    double *bound;   /* cps + 1 lower bounds, bound[cps] = L */
} bgrid;

static void bgrid_init(bgrid *g, double L, double cutoff) {
    int cps = (int)(L / cutoff);
    if (cps < 3) { g->cps = 0; g->side = L; g->bound = NULL; return; }
    g->cps = cps;
    g->side = L / cps;
    g->bound = (double *)malloc(sizeof(double) * (cps + 1));
    for (int c = 0; c < cps; c++) {
        double lo = c * g->side;
        if (lo > 0.0) {
            while ((int)(lo / g->side) == c) lo = nextafter(lo, -INFINITY);
            while ((int)(lo / g->side) < c) lo = nextafter(lo, INFINITY);
        }
        g->bound[c] = lo;
    }
    g->bound[cps] = L;
}

static int bcell1(const bgrid *g, double x) {
    int c = (int)(x / g->side);
    return c < g->cps ? c : g->cps - 1;
}

static double bmin_image(double s, double L) {
    if (s >= 0.5 * L) return s - L;
    if (s < -0.5 * L) return s + L;
    return s;
}

typedef struct {
    double d; int rank, j, f;
} bkey;
static int bkey_lt(bkey a, bkey b) {
    if (a.d != b.d) return a.d < b.d;
    if (a.rank != b.rank) return a.rank < b.rank;
    if (a.j != b.j) return a.j < b.j;
    return a.f < b.f;
}

int64_t ora_batched_run(const ora_batched_config *cfg, double *pos, ora_batched_state *st, int64_t legs,
                        ora_batched_event *log) {
    const int n = cfg->n;
    const double L = cfg->system_length;
    bgrid g;
    bgrid_init(&g, L, cfg->cutoff);
    const int ncell = g.cps ? g.cps * g.cps * g.cps : 0;
    int *cell_of = (int *)malloc(sizeof(int) * n);
    /* cell of every unit; only the active unit moves, so one entry changes per leg */
    if (g.cps)
        for (int i = 0; i < n; i++)
            cell_of[i] = bcell1(&g, pos[3 * i]) + g.cps * (bcell1(&g, pos[3 * i + 1]) + g.cps * bcell1(&g, pos[3 * i + 2]));
    int64_t k;
    for (k = 0; k < legs; k++) {
        const int a = st->active, d = st->direction;
        double ap[3] = {pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]};
        bkey best = {INFINITY, 3, 0, 0};
        /* candidates: every unit in the 27 nearby cells of the active cell, or all units */
        int ac[3] = {0, 0, 0};
        if (g.cps)
            for (int q = 0; q < 3; q++) ac[q] = bcell1(&g, ap[q]);
        for (int j = 0; j < n; j++) {
            if (j == a) continue;
            if (g.cps) {
                int cj = cell_of[j];
                int jc[3] = {cj % g.cps, (cj / g.cps) % g.cps, cj / (g.cps * g.cps)};
                int near = 1;
                for (int q = 0; q < 3; q++) {
                    int dd = ((jc[q] - ac[q]) % g.cps + g.cps) % g.cps;
                    if (!(dd == 0 || dd == 1 || dd == g.cps - 1)) near = 0;
                }
                if (!near) continue;
            }
            double s[3];
            for (int q = 0; q < 3; q++) s[q] = bmin_image(pos[3 * j + q] - ap[q], L);
            uint32_t ctr[4] = {(uint32_t)st->leg, (uint32_t)(st->leg >> 32), (uint32_t)j, 0u};
            uint32_t key[2] = {cfg->replica, (uint32_t)cfg->seed};
            uint32_t w[4];
            ora_philox4x32_10(ctr, key, w);
            for (int f = 0; f < cfg->n_factors; f++) {
                double u = words_to_unit(w[2 * f], w[2 * f + 1]);
                double dU = -ora_log(1.0 - u, cfg->math_mode) / cfg->beta;
                double dt = ora_ipt_displacement(3, d, 1.0, s, 1.0, 1.0, cfg->power[f], cfg->prefactor[f], cfg->cutoff,
                                                 dU, cfg->math_mode);
                bkey c = {dt, 0, j, f};
                if (bkey_lt(c, best)) best = c;
            }
        }
        double cb = INFINITY;
        int cbc = 0;
        if (g.cps) {
            cbc = bcell1(&g, ap[d]);
            cb = g.bound[cbc + 1] - ap[d];
            bkey c = {cb, 1, 0, 0};
            if (bkey_lt(c, best)) best = c;
        }
        {
            bkey c = {st->chain_left, 2, 0, 0};
            if (bkey_lt(c, best)) best = c;
        }
        ora_batched_event ev;
        memset(&ev, 0, sizeof(ev));
        ev.kind = best.rank;
        ev.active = a;
        ev.target = -1;
        ev.displacement = best.d;
        double x = ap[d];
        if (best.rank == 1) {
            x = cbc + 1 < g.cps ? g.bound[cbc + 1] : 0.0;
        } else {
            x = x + best.d;
            if (x >= L) x -= L;
        }
        pos[3 * a + d] = x;
        if (g.cps)
            cell_of[a] = bcell1(&g, pos[3 * a]) + g.cps * (bcell1(&g, pos[3 * a + 1]) + g.cps * bcell1(&g, pos[3 * a + 2]));
        st->est[1] += best.d;
        if (best.rank == 0) {
            int j = best.j;
            double s_d = bmin_image(pos[3 * j + d] - ap[d], L) - best.d;
            st->est[0] += s_d;
            st->est[2] += 1.0;
            st->est[4] += 1.0;
            st->chain_left -= best.d;
            st->active = j;
            ev.target = j;
            ev.factor = best.f;
        } else if (best.rank == 1) {
            st->est[3] += 1.0;
            st->chain_left -= best.d;
        } else {
            uint32_t ctr[4] = {(uint32_t)st->leg, (uint32_t)(st->leg >> 32), 0xFFFFFFFFu, 1u};
            uint32_t key[2] = {cfg->replica, (uint32_t)cfg->seed};
            uint32_t w[4];
            ora_philox4x32_10(ctr, key, w);
            int nk = (int)(((uint64_t)w[0] * (uint64_t)n) >> 32);
            st->est[5] += 1.0;
            st->chain_left = cfg->chain_length;
            st->direction = (d + 1) % 3;
            st->active = nk;
            ev.target = nk;
        }
        ev.x = x;
        st->leg++;
        if (log) log[k] = ev;
    }
    (void)ncell;
    free(cell_of);
    free(g.bound);
    return k;
}
