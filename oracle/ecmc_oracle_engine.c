/*
 * ecmc_oracle_engine.c -- cells, cell occupancy, taggers, schedulers and the leg loop.
 * TEST INFRASTRUCTURE ONLY (see ecmc_oracle.h).
 */
#include "ecmc_oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static int time_lt_pub(double q1, double r1, double q2, double r2) { return q1 < q2 || (q1 == q2 && r1 < r2); }

/* ===== CuboidPeriodicCells (cuboid_cells.py:54-297, cuboid_periodic_cells.py:74-207) === */
struct ora_cells {
    int dim, layers;
    int cps[3];
    long cum[3];
    double side[3];
    double L;
    int count;
    double *cmin, *cmax;     /* count * dim */
    int *nearby_start, *nearby; /* canonical insertion order of _yield_nearby_cells, deduplicated */
};

static int trunc_int(double x) { return (int)x; } /* Python int() truncates toward zero */

ora_cells *ora_cells_create(int dim, double L, const int *cells_per_side, int layers) {
    ora_cells *c = (ora_cells *)calloc(1, sizeof(ora_cells));
    c->dim = dim; c->layers = layers; c->L = L;
    for (int d = 0; d < dim; d++) {
        c->cps[d] = cells_per_side[d];
        c->side[d] = L / c->cps[d];
    }
    c->cum[0] = 1;
    for (int d = 1; d < dim; d++) c->cum[d] = c->cum[d - 1] * c->cps[d - 1];
    c->count = (int)(c->cum[dim - 1] * c->cps[dim - 1]);
    c->cmin = (double *)malloc(sizeof(double) * c->count * dim);
    c->cmax = (double *)malloc(sizeof(double) * c->count * dim);
    for (int cell = 0; cell < c->count; cell++) {
        for (int d = 0; d < dim; d++) {
            int id = (int)((cell / c->cum[d]) % c->cps[d]);
            double lo = id * c->side[d];
            double hi = (id + 1) * c->side[d];
            if (lo > 0.0) {
                while (trunc_int(lo / c->side[d]) == id) lo = nextafter(lo, -INFINITY);
                while (trunc_int(lo / c->side[d]) < id) lo = nextafter(lo, INFINITY);
            }
            while (trunc_int(hi / c->side[d]) == id) hi = nextafter(hi, INFINITY);
            while (trunc_int(hi / c->side[d]) > id) hi = nextafter(hi, -INFINITY);
            c->cmin[cell * dim + d] = lo;
            c->cmax[cell * dim + d] = hi;
        }
    }
    int span = 2 * layers + 1;
    int per = 1;
    for (int d = 0; d < dim; d++) per *= span;
    c->nearby_start = (int *)malloc(sizeof(int) * (c->count + 1));
    c->nearby = (int *)malloc(sizeof(int) * (size_t)c->count * per);
    int fill = 0;
    for (int cell = 0; cell < c->count; cell++) {
        c->nearby_start[cell] = fill;
        int id[3];
        for (int d = 0; d < dim; d++) id[d] = (int)((cell / c->cum[d]) % c->cps[d]);
        /* itertools.product order: the last dimension varies fastest */
        for (int t = 0; t < per; t++) {
            int rem = t, off[3];
            for (int d = dim - 1; d >= 0; d--) { off[d] = rem % span - layers; rem /= span; }
            long idx = 0;
            for (int d = 0; d < dim; d++) {
                int v = id[d] + off[d];
                v = ((v % c->cps[d]) + c->cps[d]) % c->cps[d];
                idx += v * c->cum[d];
            }
            int dup = 0;
            for (int q = c->nearby_start[cell]; q < fill; q++)
                if (c->nearby[q] == idx) { dup = 1; break; }
            if (!dup) c->nearby[fill++] = (int)idx;
        }
    }
    c->nearby_start[c->count] = fill;
    return c;
}

void ora_cells_destroy(ora_cells *c) {
    if (!c) return;
    free(c->cmin); free(c->cmax); free(c->nearby_start); free(c->nearby); free(c);
}
int ora_cells_count(const ora_cells *c) { return c->count; }
void ora_cells_bounds(const ora_cells *c, int cell, double *mn, double *mx) {
    for (int d = 0; d < c->dim; d++) { mn[d] = c->cmin[cell * c->dim + d]; mx[d] = c->cmax[cell * c->dim + d]; }
}
int ora_cells_position_to_cell(const ora_cells *c, const double *p) {
    long idx = 0;
    for (int d = 0; d < c->dim; d++) idx += (long)trunc_int(p[d] / c->side[d]) * c->cum[d];
    if (idx < 0 || idx >= c->count) return -1;
    return (int)idx;
}
int ora_cells_nearby(const ora_cells *c, int cell, int *out) {
    int n = c->nearby_start[cell + 1] - c->nearby_start[cell];
    memcpy(out, c->nearby + c->nearby_start[cell], sizeof(int) * n);
    return n;
}
int ora_cells_neighbor(const ora_cells *c, int cell, int direction, int positive) {
    int id = (int)((cell / c->cum[direction]) % c->cps[direction]);
    int nid = positive ? (id + 1) % c->cps[direction] : ((id - 1) % c->cps[direction] + c->cps[direction]) % c->cps[direction];
    return (int)(cell + (long)(nid - id) * c->cum[direction]);
}

/* cuboid_periodic_cells.py:156-180: the cell whose offset from the zero cell equals that of
 * `cell` from `ref` (cell centres, periodic), i.e. the per-axis identifier difference. */
static int cells_relative(const ora_cells *c, int cell, int ref) {
    int out = 0;
    for (int d = 0; d < c->dim; d++) {
        int a = (int)((cell / c->cum[d]) % c->cps[d]), b = (int)((ref / c->cum[d]) % c->cps[d]);
        out += (int)((((a - b) % c->cps[d]) + c->cps[d]) % c->cps[d] * c->cum[d]);
    }
    return out;
}

/* InnerPointEstimator.derivative_bound (inner_point_estimator.py:108-171) for every cell
 * separation that is not nearby the zero cell (cell_bounding_potential.py:100-126). */
void ora_cell_bound_table(const ora_cells *cells, const ora_ewald *ewald, double potential_prefactor,
                          double target_charge, int pps, double est_prefactor, double empirical_bound,
                          double system_length, double *out) {
    const int nc = cells->count, dim = cells->dim;
    const double half = system_length / 2.0;
    int near[27 * 27];
    const int nn = ora_cells_nearby(cells, 0, near);
    for (int c = 0; c < nc; c++) {
        int is_near = 0;
        for (int k = 0; k < nn; k++) if (near[k] == c) is_near = 1;
        for (int d = 0; d < 3; d++) { out[(c * 3 + d) * 2] = NAN; out[(c * 3 + d) * 2 + 1] = NAN; }
        if (is_near) continue;
        double lo[3], hi[3];
        for (int d = 0; d < dim; d++) {
            lo[d] = cells->cmin[(size_t)c * dim + d] - cells->cmax[d];
            hi[d] = cells->cmax[(size_t)c * dim + d] - cells->cmin[d];
        }
        for (int dir = 0; dir < dim; dir++) {
            double upper = -INFINITY, lower = INFINITY;
            for (int ix = 0; ix <= pps; ix++) {
                double px = lo[0] + (hi[0] - lo[0]) * ix / pps;
                for (int iy = 0; iy <= pps; iy++) {
                    double py = lo[1] + (hi[1] - lo[1]) * iy / pps;
                    for (int iz = 0; iz <= pps; iz++) {
                        double pz = lo[2] + (hi[2] - lo[2]) * iz / pps;
                        double sep[3] = {px, py, pz};
                        for (int d = 0; d < 3; d++) sep[d] = ora_py_mod(sep[d] + half, system_length) - half;
                        double p[3];
                        for (int d = 0; d < 3; d++) p[d] = sep[(dir + d) % 3];       /* permutation_3d */
                        double der = potential_prefactor * 1.0 * target_charge * ora_ewald_derivative(ewald, p[0], p[1], p[2]);
                        der = der * 1.0;                                                  /* times the speed */
                        if (der > upper) upper = der;
                        if (der < lower) lower = der;
                    }
                }
            }
            upper = upper > 0.0 ? upper * est_prefactor : upper / est_prefactor;
            lower = lower > 0.0 ? lower / est_prefactor : lower * est_prefactor;
            out[(c * 3 + dir) * 2] = upper < empirical_bound ? upper : empirical_bound;
            out[(c * 3 + dir) * 2 + 1] = lower > -empirical_bound ? lower : -empirical_bound;
        }
    }
}

/* Walker's table (walker.py:73-101).  small_list / large_list are stacks: items at or below
 * the mean go to the small one, in input order; each step pops one of each, stores
 * (small, alias large with rate mean - small.rate), lowers the large item's rate by that
 * amount and pushes it back by comparison with the mean.  The total is the sequential sum
 * of the rates (CPython before 3.12 and PyPy; 3.12 compensates the summation of floats). */
int ora_walker_build(ora_walker *w, int m, const int32_t *items, const double *rates) {
    memset(w, 0, sizeof(*w));
    if (m <= 0) return -1;
    double total = 0.0;
    for (int i = 0; i < m; i++) total += rates[i];
    const double mean = total / m;
    for (int i = 0; i < m; i++) if (!(rates[i] >= 0.0)) return -1;
    int32_t *sid = (int32_t *)malloc(sizeof(int32_t) * m), *lid = (int32_t *)malloc(sizeof(int32_t) * m);
    double *sr = (double *)malloc(sizeof(double) * m), *lr = (double *)malloc(sizeof(double) * m);
    w->item0 = (int32_t *)malloc(sizeof(int32_t) * m);
    w->item1 = (int32_t *)malloc(sizeof(int32_t) * m);
    w->rate0 = (double *)malloc(sizeof(double) * m);
    int ns = 0, nl = 0, k = 0, ok = 1;
    for (int i = 0; i < m; i++) {
        if (rates[i] > mean) { lid[nl] = items[i]; lr[nl++] = rates[i]; }
        else { sid[ns] = items[i]; sr[ns++] = rates[i]; }
    }
    while (ns && nl) {
        const int32_t si = sid[--ns], li = lid[--nl];
        const double srate = sr[ns];
        double lrate = lr[nl];
        w->item0[k] = si; w->rate0[k] = srate; w->item1[k] = li; k++;
        lrate -= mean - srate;
        if (lrate < mean) { sid[ns] = li; sr[ns++] = lrate; }
        else { lid[nl] = li; lr[nl++] = lrate; }
    }
    while (ns) {
        --ns;
        const double q = sr[ns] / mean;
        if (!(1 - 1e-6 < q && q < 1 + 1e-6)) ok = 0;
        w->item0[k] = sid[ns]; w->rate0[k] = mean; w->item1[k] = -1; k++;
    }
    while (nl) {
        --nl;
        const double q = lr[nl] / mean;
        if (!(1 - 1e-6 < q && q < 1 + 1e-6)) ok = 0;
        w->item0[k] = lid[nl]; w->rate0[k] = mean; w->item1[k] = -1; k++;
    }
    free(sid); free(lid); free(sr); free(lr);
    w->m = k; w->total = total; w->mean = mean;
    if (!ok || k != m) { ora_walker_free(w); return -1; }
    return k;
}
void ora_walker_free(ora_walker *w) {
    free(w->item0); free(w->item1); free(w->rate0);
    memset(w, 0, sizeof(*w));
}
/* walker.py:103-117: random.choice(table) = table[_randbelow(len)], then uniform(0, mean) */
int32_t ora_walker_sample(const ora_walker *w, ora_mt *r) {
    const int64_t i = ora_randint(r, 0, w->m - 1);
    const double u = ora_uniform(r, 0.0, w->mean);
    return u <= w->rate0[i] ? w->item0[i] : w->item1[i];
}

/* cuboid_periodic_cells.py:182-207: the cell holding the centre of `cell` shifted by the
 * lower corner of `rel` (positions corrected into [0, L)) */
int ora_cells_translate(const ora_cells *c, int cell, int rel, double L) {
    double cmin[3], cmax[3], rmin[3], rmax[3], p[3];
    ora_cells_bounds(c, cell, cmin, cmax);
    ora_cells_bounds(c, rel, rmin, rmax);
    for (int d = 0; d < c->dim; d++) p[d] = ora_py_mod((cmax[d] + cmin[d]) / 2.0 + rmin[d], L);
    return ora_cells_position_to_cell(c, p);
}

/* ===== small growable int list ========================================================= */
typedef struct { int *v; int n, cap; } ilist;
static void il_push(ilist *l, int x) {
    if (l->n == l->cap) { l->cap = l->cap ? 2 * l->cap : 8; l->v = (int *)realloc(l->v, sizeof(int) * l->cap); }
    l->v[l->n++] = x;
}
static int il_remove(ilist *l, int x) { /* list.remove: first occurrence */
    for (int i = 0; i < l->n; i++)
        if (l->v[i] == x) { memmove(l->v + i, l->v + i + 1, sizeof(int) * (l->n - i - 1)); l->n--; return 1; }
    return 0;
}

/* ===== handlers ======================================================================== */
typedef struct {
    int tagger;
    double cb_rate;         /* bounding event rate of a cell-bounding candidate or cell-veto event */
    int cb_cell;            /* active cell of a cell-bounding candidate; target cell of a cell veto */
    int pending;           /* event currently stored in the scheduler */
    double tq, tr;
    int target;            /* pair target or new active (end of chain) */
    /* in-state copy of the active unit, already time-sliced to the event time where the
     * reference slices it in send_event_time */
    double apos[3]; int aid; int adir; double aspeed; double aq, ar;
    double tpos[3];
    int cb_direction; double cb_boundary;
    double last_q, last_r;  /* end of chain: _last_committed_event_time */
    double samp_q, samp_r;  /* sampling: running _event_time */
    uint32_t heap_counter;  /* heap scheduler: _minimal_valid_counter */
} handler;

/* heap.c:81-326 restated on a plain array: entry 0 is a -inf sentinel, insert bubbles the new
 * entry up from the end, root pops stale entries (lazy deletion through the per-handler
 * counters, heap_scheduler.py:242-261) and bubbles the last entry down from the root.  `seq`
 * numbers the pushes; with first-push ties the root's time is shared by the valid entry
 * pushed first (the rule of the list scheduler, list_scheduler.py:129-156). */
typedef struct { double q, r; int h; uint32_t counter; uint64_t seq; } heap_entry;
typedef struct { heap_entry *a; int len, cap; uint64_t pushes; } ora_heap;

static void oheap_insert(ora_heap *H, double q, double r, int h, uint32_t counter) {
    if (H->len + 2 > H->cap) {
        int old = H->cap;
        H->cap = H->cap ? H->cap * 2 : 64;
        H->a = (heap_entry *)realloc(H->a, sizeof(heap_entry) * H->cap);
        if (!old) { H->a[0].q = -INFINITY; H->a[0].r = -INFINITY; H->a[0].h = -1; H->a[0].seq = 0; H->len = 1; }
    }
    int pos = H->len++;
    int par = pos >> 1;
    while (q < H->a[par].q || (q == H->a[par].q && r < H->a[par].r)) {
        H->a[pos] = H->a[par];
        pos = par; par = pos >> 1;
    }
    H->a[pos].q = q; H->a[pos].r = r; H->a[pos].h = h; H->a[pos].counter = counter; H->a[pos].seq = H->pushes++;
}
/* heap.c:209-247: the entry at a[len] moves down from pos */
static void oheap_bubble_down(ora_heap *H, int pos) {
    while (pos < H->len) {
        int cmp = H->len, ch = pos << 1;
        if (ch < H->len && (H->a[ch].q < H->a[cmp].q || (H->a[ch].q == H->a[cmp].q && H->a[ch].r < H->a[cmp].r))) cmp = ch;
        if (ch + 1 < H->len && (H->a[ch + 1].q < H->a[cmp].q || (H->a[ch + 1].q == H->a[cmp].q && H->a[ch + 1].r < H->a[cmp].r))) cmp = ch + 1;
        H->a[pos] = H->a[cmp];
        pos = cmp;
    }
}
typedef int (*oheap_stale_fn)(void *ctx, int h, uint32_t counter);
/* heap.c:263-276; returns the handler of the root entry after lazy deletion, or -1 */
static int oheap_root(ora_heap *H, oheap_stale_fn stale, void *ctx, int first_push_ties) {
    while (H->len > 1 && stale(ctx, H->a[1].h, H->a[1].counter)) {
        H->a[1] = H->a[--H->len];
        oheap_bubble_down(H, 1);
    }
    if (H->len <= 1) return -1;
    if (!first_push_ties) return H->a[1].h;
    /* every entry of the root's time lies in a subtree holding the root (children never come
     * before their parent): search it for the valid entry pushed first */
    int best = 1;
    int stack[64], top = 0;
    stack[top++] = 1;
    while (top > 0) {
        const int i = stack[--top];
        for (int c = 2 * i; c <= 2 * i + 1 && c < H->len; c++) {
            if (H->a[c].q != H->a[1].q || H->a[c].r != H->a[1].r) continue;
            if (!stale(ctx, H->a[c].h, H->a[c].counter) && H->a[c].seq < H->a[best].seq) best = c;
            if (top < 62) stack[top++] = c;
        }
    }
    return H->a[best].h;
}

struct ora_engine {
    ora_config cfg;
    int dim, n;
    double L, half_L;
    double *pos;            /* n * dim */
    double *charge;
    int active; int adir; double aspeed; double aq, ar;   /* lifting state */
    ora_cells *cells;
    ilist *occ;             /* per cell */
    int *sur_cell, *sur_id; /* surplus units in the order the reference's dict yields them */
    int nsur, capsur;
    double *cbt[ORA_MAX_TAGGERS]; /* cell-bound tables of cell-bounding and cell-veto taggers */
    ora_walker *cv[ORA_MAX_TAGGERS]; /* cell veto: Walker tables [direction * 2 + (0 upper, 1 lower)] */
    int occ_active_id, occ_active_cell;
    ora_mt rng;
    handler *h; int nh;
    ilist *running, *idle;  /* per tagger */
    int started, done, err;
    char errmsg[256];
    /* list scheduler */
    int *lst; int nlst, caplst;
    /* heap scheduler */
    ora_heap hp;
    double last_q, last_r;
    ora_ewald *ewald[ORA_MAX_TAGGERS];
    double est[8];
    int32_t *cand; int ncand, capcand;
    uint32_t chash;
    int pending_creates, last_tagger;
    int64_t rng_words;
};

/* Order-dependent hash term of candidate j with target t, summed over a leg's candidates
 * (a checksum for comparing candidate sets; it has no reference counterpart). */
static uint32_t cand_mix(uint32_t j, uint32_t t) {
    uint32_t h = (j * 0x9E3779B9u) ^ t;
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}

static void eng_fail(ora_engine *e, const char *msg) {
    if (!e->err) { e->err = 1; snprintf(e->errmsg, sizeof(e->errmsg), "%s", msg); }
}

/* --- schedulers ------------------------------------------------------------------------ */
static void sched_push(ora_engine *e, int hidx, double q, double r) {
    handler *hd = &e->h[hidx];
    hd->tq = q; hd->tr = r; hd->pending = 1;
    if (e->cfg.scheduler == ORA_SCHED_LIST) {
        if (e->nlst == e->caplst) { e->caplst = e->caplst ? 2 * e->caplst : 64; e->lst = (int *)realloc(e->lst, sizeof(int) * e->caplst); }
        e->lst[e->nlst++] = hidx;
        return;
    }
    /* heap_scheduler.py:139 drops infinite times; heap.c:156-189 insert */
    if (!time_lt_pub(q, r, INFINITY, INFINITY)) return;
    oheap_insert(&e->hp, q, r, hidx, hd->heap_counter);
}
static int engine_stale(void *ctx, int h, uint32_t counter) {
    const ora_engine *e = (const ora_engine *)ctx;
    return e->h[h].heap_counter > counter;
}
static int sched_pop_min(ora_engine *e) {
    int best = -1;
    if (e->cfg.scheduler == ORA_SCHED_LIST) {
        /* list_scheduler.py:129-156: min(..., key=time) keeps the first minimal element */
        for (int i = 0; i < e->nlst; i++) {
            handler *c = &e->h[e->lst[i]];
            if (best < 0 || time_lt_pub(c->tq, c->tr, e->h[best].tq, e->h[best].tr)) best = e->lst[i];
        }
    } else {
        /* heap.c:263-276 with heap_scheduler.py:340-362 (stale when counter moved on) */
        best = oheap_root(&e->hp, engine_stale, e, e->cfg.ties == ORA_TIES_FIRST_PUSH);
    }
    if (best < 0) { eng_fail(e, "SchedulerError: scheduler holds no events"); return -1; }
    if (time_lt_pub(e->h[best].tq, e->h[best].tr, e->last_q, e->last_r)) {
        eng_fail(e, "SchedulerError: event time decreased"); return -1;
    }
    e->last_q = e->h[best].tq; e->last_r = e->h[best].tr;
    return best;
}
static void sched_trash(ora_engine *e, int hidx) {
    if (e->cfg.scheduler == ORA_SCHED_LIST) {
        for (int i = 0; i < e->nlst; i++)
            if (e->lst[i] == hidx) { memmove(e->lst + i, e->lst + i + 1, sizeof(int) * (e->nlst - i - 1)); e->nlst--; e->h[hidx].pending = 0; return; }
        eng_fail(e, "SchedulerError: event handler not present in scheduler");
        return;
    }
    e->h[hidx].heap_counter++;
    e->h[hidx].pending = 0;
}

/* --- occupancy (single_active_cell_occupancy.py:95-203) ------------------------------- */
static int occ_bounded(const ora_engine *e) { return e->cfg.max_occupants > 0; }
/* self._surplus.setdefault(cell, []).append(id): join the cell's group, or open a new group
 * at the end (dict insertion order; a group that empties is deleted). */
static void sur_append(ora_engine *e, int cell, int id) {
    if (e->nsur == e->capsur) {
        e->capsur = e->capsur ? 2 * e->capsur : 64;
        e->sur_cell = (int *)realloc(e->sur_cell, sizeof(int) * e->capsur);
        e->sur_id = (int *)realloc(e->sur_id, sizeof(int) * e->capsur);
    }
    int at = e->nsur;
    for (int i = e->nsur - 1; i >= 0; i--) if (e->sur_cell[i] == cell) { at = i + 1; break; }
    memmove(e->sur_cell + at + 1, e->sur_cell + at, sizeof(int) * (e->nsur - at));
    memmove(e->sur_id + at + 1, e->sur_id + at, sizeof(int) * (e->nsur - at));
    e->sur_cell[at] = cell; e->sur_id[at] = id; e->nsur++;
}
static int sur_remove(ora_engine *e, int cell, int id) {
    for (int i = 0; i < e->nsur; i++)
        if (e->sur_cell[i] == cell && e->sur_id[i] == id) {
            memmove(e->sur_cell + i, e->sur_cell + i + 1, sizeof(int) * (e->nsur - i - 1));
            memmove(e->sur_id + i, e->sur_id + i + 1, sizeof(int) * (e->nsur - i - 1));
            e->nsur--;
            return 1;
        }
    return 0;
}
static void occ_insert(ora_engine *e, int cell, int id) {
    if (!occ_bounded(e) || e->occ[cell].n < e->cfg.max_occupants) il_push(&e->occ[cell], id);
    else sur_append(e, cell, id);
}
static void occ_update(ora_engine *e) {
    int id = e->active;
    if (id != e->occ_active_id) {
        if (e->occ_active_id >= 0) occ_insert(e, e->occ_active_cell, e->occ_active_id);
        e->occ_active_cell = ora_cells_position_to_cell(e->cells, e->pos + (size_t)id * e->dim);
        if (e->occ_active_cell < 0) { eng_fail(e, "position outside of the cell system"); return; }
        e->occ_active_id = id;
        if (!il_remove(&e->occ[e->occ_active_cell], id)) {
            /* the move from the surplus list into the occupant list never triggers in the
             * reference (the surplus dict never holds an empty list), so only remove */
            if (!sur_remove(e, e->occ_active_cell, id)) eng_fail(e, "active unit not found in occupancy");
        }
    } else {
        e->occ_active_cell = ora_cells_position_to_cell(e->cells, e->pos + (size_t)id * e->dim);
        if (e->occ_active_cell < 0) eng_fail(e, "position outside of the cell system");
    }
}

/* --- time slicing (event_handler/abstracts/abstracts.py:82-95) ------------------------ */
static void slice_pos(const ora_engine *e, double *p, int dir, double speed, double q0, double r0, double q1, double r1) {
    double dt = ora_time_sub(q1, r1, q0, r0);
    for (int d = 0; d < e->dim; d++) {
        double v = d == dir ? speed : 0.0;
        p[d] = ora_py_mod(p[d] + v * dt, e->L);
    }
}

static void separation_vector(const ora_engine *e, const double *a, const double *t, double *s) {
    for (int d = 0; d < e->dim; d++)
        s[d] = ora_py_mod((t[d] - a[d]) + e->half_L, e->L) - e->half_L;
}

/* --- pair event handlers --------------------------------------------------------------- */
static double pair_displacement(ora_engine *e, const ora_tagger *tg, int tidx, const double *s, int a, int t) {
    int mm = e->cfg.math_mode;
    int dir = e->adir; double speed = e->aspeed;
    double c1 = tg->use_charge ? e->charge[a] : 1.0;
    double c2 = tg->use_charge ? e->charge[t] : 1.0;
    if (tg->event_handler == ORA_EH_TWO_LEAF) {
        switch (tg->potential) {
        case ORA_POT_HARD_SPHERE: {
            double v[3] = {0, 0, 0};
            v[dir] = speed;
            return ora_hs_displacement(e->dim, v, s, tg->par[0]);
        }
        case ORA_POT_HARD_DIPOLE: {
            double v[3] = {0, 0, 0};
            v[dir] = speed;
            return ora_hd_displacement(e->dim, v, s, tg->par[0], tg->par[1]);
        }
        case ORA_POT_INVERSE_POWER: {
            double dU = ora_expovariate(&e->rng, e->cfg.beta, mm);
            return ora_ip_displacement(e->dim, dir, speed, s, c1, c2, tg->par[0], tg->par[1], dU, mm);
        }
        case ORA_POT_INVERSE_POWER_TRUNC: {
            double dU = ora_expovariate(&e->rng, e->cfg.beta, mm);
            return ora_ipt_displacement(e->dim, dir, speed, s, c1, c2, tg->par[0], tg->par[1], tg->par[2], dU, mm);
        }
        case ORA_POT_LENNARD_JONES: {
            double dU = ora_expovariate(&e->rng, e->cfg.beta, mm);
            return ora_lj_displacement(e->dim, dir, speed, s, tg->par[0], tg->par[1], dU, mm);
        }
        case ORA_POT_DISPLACED_EVEN_POWER: {
            double dU = ora_expovariate(&e->rng, e->cfg.beta, mm);
            return ora_dep_displacement(e->dim, dir, speed, s, tg->par[0], (int)tg->par[1], tg->par[2], dU, mm);
        }
        default: eng_fail(e, "unsupported potential for TwoLeafUnitEventHandler"); return INFINITY;
        }
    }
    if (tg->event_handler == ORA_EH_CELL_BOUNDING) {
        /* two_leaf_unit_cell_bounding_potential_event_handler.py:159-179 with
         * cell_bounding_potential.py:161-221 */
        double dU = ora_expovariate(&e->rng, e->cfg.beta, mm);
        handler *hd = &e->h[tidx];
        int tcell = ora_cells_position_to_cell(e->cells, e->pos + (size_t)t * e->dim);
        int rel = cells_relative(e->cells, tcell, e->occ_active_cell);
        const double *row = e->cbt[hd->tagger] + (size_t)(rel * 3 + dir) * 2;
        double rate;
        if (tg->use_charge) {
            double cf = c1 * c2 / tg->bpar[2];                  /* charge_correction_factor */
            rate = cf > 0.0 ? row[0] * cf : row[1] * cf;
        } else {
            rate = row[0];
        }
        hd->cb_rate = rate;
        hd->cb_cell = e->occ_active_cell;
        double disp = rate > 0 ? dU / rate : INFINITY;
        return disp / speed;
    }
    if (tg->event_handler == ORA_EH_TWO_LEAF_BOUNDING && tg->bounding_potential == ORA_POT_IPCB) {
        /* two_leaf_unit_bounding_potential_event_handler.py:138-143 */
        double p[3];
        for (int d = 0; d < 3; d++) p[d] = s[(dir + d) % 3];    /* vectors.permutation_3d */
        double pp = tg->bpar[0] * c1 * c2;
        double dU = ora_expovariate(&e->rng, e->cfg.beta, mm);
        return ora_ipcb_displacement(pp, p[0], p[1], p[2], dU, e->L) / speed;
    }
    eng_fail(e, "unsupported event handler");
    return INFINITY;
}

/* --- leg helpers ----------------------------------------------------------------------- */
static int pool_pop(ora_engine *e, int t) {
    ilist *idle = &e->idle[t];
    if (idle->n == 0) { eng_fail(e, "TagActivatorError: not-running event handler list is empty"); return -1; }
    int hidx = idle->v[--idle->n];
    il_push(&e->running[t], hidx);
    return hidx;
}

static void create_tagger(ora_engine *e, int t) {
    ora_tagger *tg = &e->cfg.taggers[t];
    int a = e->active;
    switch (tg->kind) {
    case ORA_TAG_PAIR_CELLS:
    case ORA_TAG_PAIR_ALL:
    case ORA_TAG_SURPLUS:
    case ORA_TAG_CELL_BOUNDING:
    case ORA_TAG_FACTOR_MAP: {
        if (a < 0) return;
        /* the candidate targets in the order the tagger yields them:
         * excluded_cells_tagger.py:129-132 (nearby cells, occupants in list order),
         * surplus_cells_tagger.py:128-131 (the surplus dict's order),
         * cell_bounding_potential_tagger.py:150-155 (cells in yield_cells() order that are
         * occupied and not nearby), or every other unit in identifier order */
        const int P = e->cfg.leaves_per_root > 1 ? e->cfg.leaves_per_root : 1;
        int *targets = (int *)malloc(sizeof(int) * ((size_t)(e->n + 1) * ORA_MAX_FACTOR_MAP));
        int nt = 0;
        int cells[27 * 27];
        int ncells = 0;
        if (tg->kind == ORA_TAG_PAIR_CELLS || tg->kind == ORA_TAG_CELL_BOUNDING) {
            if (e->occ_active_cell < 0) { free(targets); return; }
            ncells = ora_cells_nearby(e->cells, e->occ_active_cell, cells);
        }
        if (tg->kind == ORA_TAG_FACTOR_MAP) {
            /* factor_type_maps.py:325-356: local factors of the active leaf's own object, or for
             * every other root in identifier order the map's entries of the active leaf number */
            const int m = a / P, k = a % P;
            if (tg->fmap_local) {
                for (int x = 0; x < tg->fmap_n[k]; x++) targets[nt++] = m * P + tg->fmap_target[k][x];
            } else {
                for (int other = 0; other < e->n / P; other++) {
                    if (other == m) continue;
                    for (int x = 0; x < tg->fmap_n[k]; x++) {
                        const int tl = tg->fmap_target[k][x];
                        targets[nt++] = tl < P ? m * P + tl : other * P + (tl - P);
                    }
                }
            }
        } else if (tg->kind == ORA_TAG_PAIR_CELLS) {
            for (int c = 0; c < ncells; c++)
                for (int o = 0; o < e->occ[cells[c]].n; o++) targets[nt++] = e->occ[cells[c]].v[o];
        } else if (tg->kind == ORA_TAG_PAIR_ALL) {
            for (int o = 0; o < e->n; o++) targets[nt++] = o;
        } else if (tg->kind == ORA_TAG_SURPLUS) {
            for (int o = 0; o < e->nsur; o++) targets[nt++] = e->sur_id[o];
        } else {
            int nc = ora_cells_count(e->cells);
            for (int c = 0; c < nc; c++) {
                if (e->occ[c].n == 0) continue;
                int near = 0;
                for (int k = 0; k < ncells; k++) if (cells[k] == c) near = 1;
                if (near) continue;
                if (e->occ[c].n != 1) { free(targets); eng_fail(e, "cell bounding needs at most one occupant per cell"); return; }
                targets[nt++] = e->occ[c].v[0];
            }
        }
        for (int k = 0; k < nt; k++) {
            int tid = targets[k];
            if (tid == a) continue;
            int hidx = pool_pop(e, t);
            if (hidx < 0) { free(targets); return; }
            handler *hd = &e->h[hidx];
            hd->aid = a; hd->adir = e->adir; hd->aspeed = e->aspeed; hd->aq = e->aq; hd->ar = e->ar;
            memcpy(hd->apos, e->pos + (size_t)a * e->dim, sizeof(double) * e->dim);
            memcpy(hd->tpos, e->pos + (size_t)tid * e->dim, sizeof(double) * e->dim);
            hd->target = tid;
            double s[3];
            separation_vector(e, hd->apos, hd->tpos, s);
            double dt = pair_displacement(e, tg, hidx, s, a, tid);
            double q, r;
            ora_time_add(e->aq, e->ar, dt, &q, &r);
            /* send_event_time time-slices the in-state copies to the event time */
            if (!isinf(q)) slice_pos(e, hd->apos, e->adir, e->aspeed, e->aq, e->ar, q, r);
            hd->aq = q; hd->ar = r;
            sched_push(e, hidx, q, r);
            if (e->ncand == e->capcand) { e->capcand = e->capcand ? 2 * e->capcand : 1024; e->cand = (int32_t *)realloc(e->cand, sizeof(int32_t) * e->capcand); }
            e->chash += cand_mix((uint32_t)e->ncand, (uint32_t)tid);
            e->cand[e->ncand++] = tid;
        }
        free(targets);
        break;
    }
    case ORA_TAG_CELL_VETO: {
        /* cell_veto_tagger.py:150-155 yields the active unit once; send_event_time
         * (cell_veto_event_handler.py:190-238): Walker table by the sign of the charge
         * correction factor, a sampled relative cell, then expovariate / (total rate * speed) */
        if (a < 0) return;
        int hidx = pool_pop(e, t);
        if (hidx < 0) return;
        handler *hd = &e->h[hidx];
        hd->aid = a; hd->adir = e->adir; hd->aspeed = e->aspeed; hd->aq = e->aq; hd->ar = e->ar;
        memcpy(hd->apos, e->pos + (size_t)a * e->dim, sizeof(double) * e->dim);
        hd->target = -1;
        const int acell = ora_cells_position_to_cell(e->cells, hd->apos);
        if (acell < 0) { eng_fail(e, "position outside of the cell system"); return; }
        double cf = (tg->use_charge ? e->charge[a] : 1.0) * 1.0 / tg->bpar[2];
        int b = 0;
        if (!(cf > 0.0)) { cf *= -1.0; b = 1; }
        const ora_walker *wk = &e->cv[t][e->adir * 2 + b];
        const double total = wk->total * cf;
        const int rel = ora_walker_sample(wk, &e->rng);
        const double row = e->cbt[t][(size_t)(rel * 3 + e->adir) * 2 + b];
        hd->cb_rate = (b ? -row : row) * cf;
        if (!(hd->cb_rate > 0.0)) { eng_fail(e, "cell veto: bounding event rate not positive"); return; }
        hd->cb_cell = ora_cells_translate(e->cells, acell, rel, e->L);
        const double dU = ora_expovariate(&e->rng, e->cfg.beta, e->cfg.math_mode);
        const double dt = dU / (total * e->aspeed);
        double q, r;
        ora_time_add(e->aq, e->ar, dt, &q, &r);
        if (!isinf(q)) slice_pos(e, hd->apos, e->adir, e->aspeed, e->aq, e->ar, q, r);
        hd->aq = q; hd->ar = r;
        sched_push(e, hidx, q, r);
        break;
    }
    case ORA_TAG_CELL_BOUNDARY: {
        if (a < 0) return;
        int hidx = pool_pop(e, t);
        if (hidx < 0) return;
        handler *hd = &e->h[hidx];
        hd->aid = a; hd->adir = e->adir; hd->aspeed = e->aspeed; hd->aq = e->aq; hd->ar = e->ar;
        memcpy(hd->apos, e->pos + (size_t)a * e->dim, sizeof(double) * e->dim);
        /* cell_boundary_event_handler.py:98-156 (velocity has a single positive component) */
        int cell = ora_cells_position_to_cell(e->cells, hd->apos);
        if (cell < 0) { eng_fail(e, "position outside of the cell system"); return; }
        int nb = ora_cells_neighbor(e->cells, cell, e->adir, 1);
        double mn[3], mx[3];
        ora_cells_bounds(e->cells, nb, mn, mx);
        double boundary = mn[e->adir];
        double sep = boundary - hd->apos[e->adir];
        if (sep < 0.0) sep = sep + e->L;
        double dt = sep / e->aspeed;
        hd->cb_boundary = boundary; hd->cb_direction = e->adir;
        double q, r;
        ora_time_add(e->aq, e->ar, dt, &q, &r);
        sched_push(e, hidx, q, r);
        break;
    }
    case ORA_TAG_SAMPLING: {
        int hidx = pool_pop(e, t);
        if (hidx < 0) return;
        handler *hd = &e->h[hidx];
        double q, r;
        ora_time_add(hd->samp_q, hd->samp_r, tg->value, &q, &r);
        hd->samp_q = q; hd->samp_r = r;
        sched_push(e, hidx, q, r);
        break;
    }
    case ORA_TAG_END_OF_CHAIN: {
        if (a < 0) return;
        int hidx = pool_pop(e, t);
        if (hidx < 0) return;
        handler *hd = &e->h[hidx];
        double since = ora_time_sub(e->aq, e->ar, hd->last_q, hd->last_r);
        if (!(0.0 <= since && since <= tg->value)) { eng_fail(e, "end-of-chain assertion on chain time failed"); return; }
        double chain = ora_time_sub(hd->last_q, hd->last_r, e->aq, e->ar) + tg->value;
        double q, r;
        ora_time_add(e->aq, e->ar, chain, &q, &r);
        /* ..._end_of_chain_event_handler.py:213-237: a random root, and with composite objects
         * then a random leaf of it */
        {
            const int P = e->cfg.leaves_per_root > 1 ? e->cfg.leaves_per_root : 1;
            const int root = (int)ora_randint(&e->rng, 0, e->n / P - 1);
            hd->target = P > 1 ? root * P + (int)ora_randint(&e->rng, 0, P - 1) : root;
        }
        sched_push(e, hidx, q, r);
        break;
    }
    case ORA_TAG_END_OF_RUN: {
        int hidx = pool_pop(e, t);
        if (hidx < 0) return;
        double q, r;
        ora_py_divmod(tg->value, 1.0, &q, &r);   /* Time.from_float */
        sched_push(e, hidx, q, r);
        break;
    }
    case ORA_TAG_START_OF_RUN: {
        int hidx = pool_pop(e, t);
        if (hidx < 0) return;
        sched_push(e, hidx, 0.0, 0.0);
        break;
    }
    }
}

static void commit_active_slice(ora_engine *e, double q, double r) {
    double *p = e->pos + (size_t)e->active * e->dim;
    slice_pos(e, p, e->adir, e->aspeed, e->aq, e->ar, q, r);
    e->aq = q; e->ar = r;
}

/* out-state + commit; fills the event record */
static int out_state(ora_engine *e, int hidx, ora_event *ev) {
    handler *hd = &e->h[hidx];
    ora_tagger *tg = &e->cfg.taggers[hd->tagger];
    ev->tagger = hd->tagger; ev->flags = 0; ev->active = e->active; ev->target = -1;
    ev->tq = hd->tq; ev->tr = hd->tr;
    int a = e->active;
    switch (tg->kind) {
    case ORA_TAG_PAIR_CELLS:
    case ORA_TAG_PAIR_ALL:
    case ORA_TAG_SURPLUS:
    case ORA_TAG_CELL_BOUNDING:
    case ORA_TAG_FACTOR_MAP: {
        ev->target = hd->target;
        int lift = 1;
        if (tg->event_handler == ORA_EH_CELL_BOUNDING) {
            /* two_leaf_unit_cell_bounding_potential_event_handler.py:181-211: confirm the
             * active cell, bound = the candidate's rate, then thin against the real derivative */
            if (ora_cells_position_to_cell(e->cells, hd->apos) != hd->cb_cell) {
                eng_fail(e, "cell-bounding event after the active unit left its cell");
                return -1;
            }
            double s[3], p[3];
            separation_vector(e, hd->apos, hd->tpos, s);
            for (int d = 0; d < 3; d++) p[d] = s[(hd->adir + d) % 3];
            double c1 = tg->use_charge ? e->charge[hd->aid] : 1.0;
            double c2 = tg->use_charge ? e->charge[hd->target] : 1.0;
            double bound = hd->cb_rate * hd->aspeed;
            double real = tg->par[3] * c1 * c2 * ora_ewald_derivative(e->ewald[hd->tagger], p[0], p[1], p[2]) * hd->aspeed;
            lift = 0;
            if (real > 0) {
                double u = ora_uniform(&e->rng, 0, bound);
                if (u < real) lift = 1;
            }
        } else if (tg->event_handler == ORA_EH_TWO_LEAF_BOUNDING) {
            /* two_leaf_unit_bounding_potential_event_handler.py:148-168 and
             * event_handler_with_bounding_potential.py:75-101 */
            double s[3], p[3];
            separation_vector(e, hd->apos, hd->tpos, s);
            for (int d = 0; d < 3; d++) p[d] = s[(hd->adir + d) % 3];
            double c1 = tg->use_charge ? e->charge[hd->aid] : 1.0;
            double c2 = tg->use_charge ? e->charge[hd->target] : 1.0;
            double bound = ora_ipcb_derivative(tg->bpar[0] * c1 * c2, p[0], p[1], p[2]) * hd->aspeed;
            double real = tg->par[3] * c1 * c2 * ora_ewald_derivative(e->ewald[hd->tagger], p[0], p[1], p[2]) * hd->aspeed;
            lift = 0;
            if (real > 0) {
                double u = ora_uniform(&e->rng, 0, bound);
                    if (u < real) lift = 1;
            }
        }
        memcpy(e->pos + (size_t)a * e->dim, hd->apos, sizeof(double) * e->dim);
        e->aq = hd->aq; e->ar = hd->ar;
        if (!lift) e->est[6] += 1.0;
        if (lift) {
            double sep[3];
            separation_vector(e, hd->apos, hd->tpos, sep);
            e->est[0] += sep[hd->adir];
            e->est[4] += 1.0;
            e->active = hd->target;       /* velocity and time stamp move to the target */
            ev->flags |= 1;
        }
        e->est[2] += 1.0;
        break;
    }
    case ORA_TAG_CELL_VETO: {
        /* mediator.py:265-292 hands over the occupant of the sampled cell (None when empty);
         * leaf_unit_cell_veto_event_handler.py:127-149 with
         * event_handler_with_bounding_potential.py:75-101: thinning of the stored rate */
        const ilist *o = &e->occ[hd->cb_cell];
        if (o->n > 1) { eng_fail(e, "cell veto: the sampled cell holds more than one unit"); return -1; }
        int lift = 0;
        double sep[3] = {0, 0, 0};
        if (o->n == 1) {
            const int tid = o->v[0];
            ev->target = tid;
            double p[3];
            separation_vector(e, hd->apos, e->pos + (size_t)tid * e->dim, sep);
            for (int d = 0; d < 3; d++) p[d] = sep[(hd->adir + d) % 3];
            double c1 = tg->use_charge ? e->charge[hd->aid] : 1.0;
            double c2 = tg->use_charge ? e->charge[tid] : 1.0;
            double real = tg->par[3] * c1 * c2 * ora_ewald_derivative(e->ewald[hd->tagger], p[0], p[1], p[2]) * hd->aspeed;
            if (real > 0) {
                double u = ora_uniform(&e->rng, 0, hd->cb_rate);
                if (u < real) lift = 1;
            }
        }
        memcpy(e->pos + (size_t)a * e->dim, hd->apos, sizeof(double) * e->dim);
        e->aq = hd->aq; e->ar = hd->ar;
        if (lift) {
            e->est[0] += sep[hd->adir];
            e->est[4] += 1.0;
            e->active = ev->target;
            ev->flags |= 1;
        } else {
            e->est[6] += 1.0;
        }
        e->est[2] += 1.0;
        break;
    }
    case ORA_TAG_CELL_BOUNDARY: {
        double *p = e->pos + (size_t)a * e->dim;
        memcpy(p, hd->apos, sizeof(double) * e->dim);
        slice_pos(e, p, hd->adir, hd->aspeed, hd->aq, hd->ar, hd->tq, hd->tr);
        p[hd->cb_direction] = hd->cb_boundary;
        e->aq = hd->tq; e->ar = hd->tr;
        e->est[3] += 1.0;
        break;
    }
    case ORA_TAG_SAMPLING:
        commit_active_slice(e, hd->tq, hd->tr);
        break;
    case ORA_TAG_END_OF_RUN:
        commit_active_slice(e, hd->tq, hd->tr);
        e->done = 1;
        break;
    case ORA_TAG_END_OF_CHAIN: {
        commit_active_slice(e, hd->tq, hd->tr);
        hd->last_q = hd->tq; hd->last_r = hd->tr;
        ev->target = hd->target;
        e->est[1] += tg->value * e->aspeed;
        e->est[5] += 1.0;
        e->adir = (e->adir + 1) % e->dim;
        if (hd->target != a) ev->flags |= 1;
        e->active = hd->target;
        break;
    }
    case ORA_TAG_START_OF_RUN:
        e->active = tg->ivalue2;
        e->adir = tg->ivalue;
        e->aspeed = tg->value;
        e->aq = 0.0; e->ar = 0.0;
        break;
    }
    memset(ev->pos, 0, sizeof(ev->pos));
    if (a >= 0) memcpy(ev->pos, e->pos + (size_t)a * e->dim, sizeof(double) * e->dim);
    return 0;
}

/* ===== engine API ======================================================================= */
ora_engine *ora_engine_create(const ora_config *cfg, const double *positions, const double *charges) {
    ora_engine *e = (ora_engine *)calloc(1, sizeof(ora_engine));
    e->cfg = *cfg;
    if (cfg->n_taggers < 1 || cfg->n_taggers > ORA_MAX_TAGGERS) {
        e->cfg.n_taggers = 0;
        eng_fail(e, "number of taggers out of range");
        return e;
    }
    e->dim = cfg->dim; e->n = cfg->n; e->L = cfg->system_length; e->half_L = cfg->system_length / 2.0;
    e->pos = (double *)malloc(sizeof(double) * e->n * e->dim);
    memcpy(e->pos, positions, sizeof(double) * e->n * e->dim);
    e->charge = (double *)malloc(sizeof(double) * e->n);
    for (int i = 0; i < e->n; i++) e->charge[i] = charges ? charges[i] : 1.0;
    e->active = -1; e->adir = 0; e->aspeed = 1.0;
    e->last_q = -INFINITY; e->last_r = -INFINITY;
    int cps[3] = {cfg->cells_per_side[0], cfg->cells_per_side[1], cfg->cells_per_side[2]};
    for (int d = 1; d < e->dim; d++) if (cps[d] <= 0) cps[d] = cps[0];
    e->cells = ora_cells_create(e->dim, e->L, cps, cfg->neighbor_layers);
    int nc = ora_cells_count(e->cells);
    e->occ = (ilist *)calloc(nc, sizeof(ilist));
    e->occ_active_id = -1; e->occ_active_cell = -1;
    for (int i = 0; i < e->n; i++) {
        int c = ora_cells_position_to_cell(e->cells, e->pos + (size_t)i * e->dim);
        if (c < 0) { eng_fail(e, "initial position outside of the cell system"); return e; }
        occ_insert(e, c, i);
    }
    ora_mt_seed(&e->rng, cfg->seed);
    for (int32_t k = 0; k < cfg->rng_skip; k++) ora_mt_u32(&e->rng);
    int total = 0;
    for (int t = 0; t < cfg->n_taggers; t++) total += cfg->taggers[t].n_handlers > 0 ? cfg->taggers[t].n_handlers : 1;
    e->h = (handler *)calloc(total, sizeof(handler));
    e->nh = total;
    e->running = (ilist *)calloc(cfg->n_taggers, sizeof(ilist));
    e->idle = (ilist *)calloc(cfg->n_taggers, sizeof(ilist));
    int k = 0;
    for (int t = 0; t < cfg->n_taggers; t++) {
        int nhh = cfg->taggers[t].n_handlers > 0 ? cfg->taggers[t].n_handlers : 1;
        for (int j = 0; j < nhh; j++) {
            e->h[k].tagger = t;
            if (cfg->taggers[t].kind == ORA_TAG_SAMPLING && cfg->taggers[t].ivalue)
                ora_py_divmod(-cfg->taggers[t].value, 1.0, &e->h[k].samp_q, &e->h[k].samp_r);
            il_push(&e->idle[t], k++);
        }
        if (cfg->taggers[t].potential == ORA_POT_MERGED_IMAGE_COULOMB)
            e->ewald[t] = ora_ewald_create((int)cfg->taggers[t].par[1], (int)cfg->taggers[t].par[2],
                                           cfg->taggers[t].par[0], e->L);
        if (cfg->taggers[t].event_handler == ORA_EH_CELL_BOUNDING || cfg->taggers[t].event_handler == ORA_EH_CELL_VETO) {
            const ora_tagger *tg = &cfg->taggers[t];
            if (e->dim != 3 || !e->ewald[t] || cfg->max_occupants != 1) {
                eng_fail(e, "cell bounding needs 3-D merged-image Coulomb and one occupant per cell");
                return e;
            }
            if ((tg->kind == ORA_TAG_CELL_VETO) != (tg->event_handler == ORA_EH_CELL_VETO)) {
                eng_fail(e, "the cell-veto handler runs under the cell-veto tagger only");
                return e;
            }
            e->cbt[t] = (double *)malloc(sizeof(double) * (size_t)nc * 6);
            ora_cell_bound_table(e->cells, e->ewald[t], tg->par[3], tg->bpar[2], (int)tg->bpar[1], tg->bpar[0],
                                 tg->bpar[3], e->L, e->cbt[t]);
        }
        if (cfg->taggers[t].event_handler == ORA_EH_CELL_VETO) {
            /* cell_veto_event_handler.py:147-159: per direction, the non-nearby cells in
             * yield_cells() order with rates max(upper, 0) and max(-lower, 0) */
            int32_t *items = (int32_t *)malloc(sizeof(int32_t) * nc);
            double *ru = (double *)malloc(sizeof(double) * nc), *rl = (double *)malloc(sizeof(double) * nc);
            e->cv[t] = (ora_walker *)calloc(6, sizeof(ora_walker));
            for (int d = 0; d < 3 && !e->err; d++) {
                int m = 0;
                for (int c = 0; c < nc; c++) {
                    const double up = e->cbt[t][(size_t)(c * 3 + d) * 2], lo = e->cbt[t][(size_t)(c * 3 + d) * 2 + 1];
                    if (isnan(up)) continue;                      /* nearby the zero cell */
                    items[m] = c;
                    ru[m] = 0.0 > up ? 0.0 : up;
                    rl[m] = 0.0 > -lo ? 0.0 : -lo;
                    m++;
                }
                if (ora_walker_build(&e->cv[t][d * 2], m, items, ru) < 0 ||
                    ora_walker_build(&e->cv[t][d * 2 + 1], m, items, rl) < 0)
                    eng_fail(e, "cell veto: Walker table failed");
            }
            free(items); free(ru); free(rl);
            if (e->err) return e;
        }
    }
    return e;
}

void ora_engine_destroy(ora_engine *e) {
    if (!e) return;
    int nc = ora_cells_count(e->cells);
    for (int c = 0; c < nc; c++) free(e->occ[c].v);
    for (int t = 0; t < e->cfg.n_taggers; t++) {
        free(e->running[t].v); free(e->idle[t].v); ora_ewald_destroy(e->ewald[t]); free(e->cbt[t]);
        if (e->cv[t]) { for (int k = 0; k < 6; k++) ora_walker_free(&e->cv[t][k]); free(e->cv[t]); }
    }
    free(e->occ); free(e->sur_cell); free(e->sur_id); free(e->running); free(e->idle);
    ora_cells_destroy(e->cells);
    free(e->pos); free(e->charge); free(e->h); free(e->lst); free(e->hp.a); free(e->cand);
    free(e);
}

int64_t ora_engine_run(ora_engine *e, int64_t max_events, ora_event *log) {
    int64_t k = 0;
    if (e->err) return -1;
    while (k < max_events && !e->done) {
        if (!e->started) {
            /* tag_activator.py:166-226: the start-of-run tagger alone */
            int sor = -1;
            for (int t = 0; t < e->cfg.n_taggers; t++) if (e->cfg.taggers[t].kind == ORA_TAG_START_OF_RUN) sor = t;
            if (sor < 0) { eng_fail(e, "ConfigurationError: a StartOfRunEventHandler is required"); return -1; }
            e->ncand = 0;
            e->chash = 0;
            create_tagger(e, sor);
            e->started = 1;
        } else if (e->pending_creates) {
            /* tag_activator.py:228-289: occupancy update, then the creates of the tagger
             * that committed the preceding event, in INI order */
            occ_update(e);
            if (e->err) return -1;
            e->ncand = 0;
            e->chash = 0;
            ora_tagger *pt = &e->cfg.taggers[e->last_tagger];
            for (int i = 0; i < pt->n_create; i++) {
                create_tagger(e, pt->create[i]);
                if (e->err) return -1;
            }
            e->pending_creates = 0;
        }
        int win = sched_pop_min(e);
        if (win < 0) return -1;
        ora_event ev;
        out_state(e, win, &ev);
        ev.ncand = e->ncand;
        ev.chash = e->chash;
        int wt = e->h[win].tagger;
        /* trash (tag_activator.py:291-316) */
        ora_tagger *tg = &e->cfg.taggers[wt];
        for (int i = 0; i < tg->n_trash; i++) {
            int t = tg->trash[i];
            for (int j = 0; j < e->running[t].n; j++) {
                int hidx = e->running[t].v[j];
                sched_trash(e, hidx);
                il_push(&e->idle[t], hidx);
            }
            e->running[t].n = 0;
        }
        if (e->err) return -1;
        if (log) log[k] = ev;
        k++;
        e->pending_creates = 1;
        e->last_tagger = wt;
    }
    return k;
}

int ora_engine_done(const ora_engine *e) { return e->done; }
int ora_engine_error(const ora_engine *e, char *buf, int buflen) {
    if (buf && buflen > 0) snprintf(buf, buflen, "%s", e->errmsg);
    return e->err;
}
void ora_engine_positions(const ora_engine *e, double *out) { memcpy(out, e->pos, sizeof(double) * e->n * e->dim); }
int32_t ora_engine_active(const ora_engine *e, int32_t *direction, double *tq, double *tr) {
    if (direction) *direction = e->adir;
    if (tq) *tq = e->aq;
    if (tr) *tr = e->ar;
    return e->active;
}
void ora_engine_estimators(const ora_engine *e, double *out8) { memcpy(out8, e->est, sizeof(e->est)); }
int32_t ora_engine_last_candidates(const ora_engine *e, int32_t *out, int32_t cap) {
    int n = e->ncand < cap ? e->ncand : cap;
    if (out) memcpy(out, e->cand, sizeof(int32_t) * n);
    return e->ncand;
}
int64_t ora_engine_rng_words(const ora_engine *e) { return (int64_t)e->rng.words; }

/* ===== configuration helpers ============================================================ */
static void tg_lists(ora_tagger *t, int nc, const int *c, int ntr, const int *tr) {
    t->n_create = nc; for (int i = 0; i < nc; i++) t->create[i] = c[i];
    t->n_trash = ntr; for (int i = 0; i < ntr; i++) t->trash[i] = tr[i];
}

/* Wiring of config_files/hard_disk_dipoles/hard_disk_dipoles_cells.ini without the dipole
 * and sampling taggers: pair(0), cell_boundary(1), end_of_chain(2), end_of_run(3), start_of_run(4). */
void ora_config_hard_disks(ora_config *c, int n, double L, double radius, int cps, double chain_time,
                           double end_time, uint64_t seed) {
    memset(c, 0, sizeof(*c));
    c->dim = 2; c->n = n; c->system_length = L; c->beta = 1.0;
    c->cells_per_side[0] = c->cells_per_side[1] = cps; c->neighbor_layers = 1;
    c->max_occupants = -1; c->scheduler = ORA_SCHED_LIST; c->math_mode = ORA_MATH_LIBM; c->seed = seed;
    c->n_taggers = 5;
    int pc[2] = {0, 1}, eoc[3] = {2, 0, 1}, all[5] = {2, 0, 1, 4, 3}, eor[1] = {3}, sorc[4] = {0, 1, 2, 3}, sort[1] = {4};
    ora_tagger *t = c->taggers;
    t[0].kind = ORA_TAG_PAIR_CELLS; t[0].event_handler = ORA_EH_TWO_LEAF; t[0].potential = ORA_POT_HARD_SPHERE;
    t[0].par[0] = radius; t[0].n_handlers = n; tg_lists(&t[0], 2, pc, 2, pc);
    t[1].kind = ORA_TAG_CELL_BOUNDARY; t[1].n_handlers = 1; tg_lists(&t[1], 2, pc, 2, pc);
    t[2].kind = ORA_TAG_END_OF_CHAIN; t[2].n_handlers = 1; t[2].value = chain_time; tg_lists(&t[2], 3, eoc, 3, eoc);
    t[3].kind = ORA_TAG_END_OF_RUN; t[3].n_handlers = 1; t[3].value = end_time; tg_lists(&t[3], 1, eor, 5, all);
    t[4].kind = ORA_TAG_START_OF_RUN; t[4].n_handlers = 1; t[4].value = 1.0; t[4].ivalue = 0; t[4].ivalue2 = 0;
    tg_lists(&t[4], 4, sorc, 1, sort);
}

/* Inverse-power factors (e.g. LJ 12-6 as p=12,k=4 and p=6,k=-4), each an ExcludedCellsTagger
 * with a TwoLeafUnitEventHandler; pair taggers 0..F-1, cell_boundary F, end_of_chain F+1,
 * end_of_run F+2, start_of_run F+3.  cutoff > 0 selects the truncated-shifted variant. */
void ora_config_inverse_power_pairs(ora_config *c, int dim, int n, double L, double beta, int cps, int nf,
                                    const double *powers, const double *prefactors, double cutoff, double chain_time,
                                    double end_time, uint64_t seed, int math_mode) {
    memset(c, 0, sizeof(*c));
    c->dim = dim; c->n = n; c->system_length = L; c->beta = beta;
    for (int d = 0; d < 3; d++) c->cells_per_side[d] = cps;
    c->neighbor_layers = 1; c->max_occupants = -1; c->scheduler = ORA_SCHED_HEAP; c->math_mode = math_mode; c->seed = seed;
    c->n_taggers = nf + 4;
    int cb = nf, eoc = nf + 1, eor = nf + 2, sor = nf + 3;
    int pc[ORA_MAX_TAGGERS], np = 0;
    for (int f = 0; f < nf; f++) pc[np++] = f;
    pc[np++] = cb;
    int eocl[ORA_MAX_TAGGERS], ne = 0;
    eocl[ne++] = eoc; for (int i = 0; i < np; i++) eocl[ne++] = pc[i];
    int all[ORA_MAX_TAGGERS], na = 0;
    for (int i = 0; i < ne; i++) all[na++] = eocl[i];
    all[na++] = sor; all[na++] = eor;
    int sorc[ORA_MAX_TAGGERS], ns = 0;
    for (int i = 0; i < np; i++) sorc[ns++] = pc[i];
    sorc[ns++] = eoc; sorc[ns++] = eor;
    int eorl[1] = {eor}, sorl[1] = {sor};
    ora_tagger *t = c->taggers;
    for (int f = 0; f < nf; f++) {
        t[f].kind = ORA_TAG_PAIR_CELLS; t[f].event_handler = ORA_EH_TWO_LEAF;
        t[f].potential = cutoff > 0 ? ORA_POT_INVERSE_POWER_TRUNC : ORA_POT_INVERSE_POWER;
        t[f].par[0] = powers[f]; t[f].par[1] = prefactors[f]; t[f].par[2] = cutoff;
        t[f].n_handlers = n; tg_lists(&t[f], np, pc, np, pc);
    }
    t[cb].kind = ORA_TAG_CELL_BOUNDARY; t[cb].n_handlers = 1; tg_lists(&t[cb], np, pc, np, pc);
    t[eoc].kind = ORA_TAG_END_OF_CHAIN; t[eoc].n_handlers = 1; t[eoc].value = chain_time; tg_lists(&t[eoc], ne, eocl, ne, eocl);
    t[eor].kind = ORA_TAG_END_OF_RUN; t[eor].n_handlers = 1; t[eor].value = end_time; tg_lists(&t[eor], 1, eorl, na, all);
    t[sor].kind = ORA_TAG_START_OF_RUN; t[sor].n_handlers = 1; t[sor].value = 1.0; t[sor].ivalue = 0; t[sor].ivalue2 = 0;
    tg_lists(&t[sor], ns, sorc, 1, sorl);
}

/* ===== CPU baseline: independent replicas on host threads ============================== */
typedef struct {
    ora_config cfg; const double *pos; const double *charge; int64_t events; int64_t done;
} thr_arg;
static void *thr_main(void *p) {
    thr_arg *a = (thr_arg *)p;
    ora_engine *e = ora_engine_create(&a->cfg, a->pos, a->charge);
    a->done = ora_engine_run(e, a->events, NULL);
    ora_engine_destroy(e);
    return NULL;
}
int64_t ora_run_threads(const ora_config *cfg, const double *positions, const double *charges, int n_threads,
                        int64_t events_per_replica, double *seconds) {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
    thr_arg *args = (thr_arg *)calloc(n_threads, sizeof(thr_arg));
    struct timespec t0, t1;
    for (int i = 0; i < n_threads; i++) {
        args[i].cfg = *cfg; args[i].cfg.seed = cfg->seed + (uint64_t)i;
        args[i].pos = positions; args[i].charge = charges; args[i].events = events_per_replica;
    }
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < n_threads; i++) pthread_create(&th[i], NULL, thr_main, &args[i]);
    int64_t total = 0;
    for (int i = 0; i < n_threads; i++) { pthread_join(th[i], NULL); total += args[i].done > 0 ? args[i].done : 0; }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds) *seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    free(th); free(args);
    return total;
}

/* ---- the heap alone (test entry) ----------------------------------------------------------
 * op[i] = 0: push (h[i], q[i], r[i]) with the handler's current counter; 1: trash h[i] (its
 * counter moves on); 2: root -> out[i] = handler of the root entry or -1.  Handlers
 * 0..n_handlers-1. */
struct replay_ctx { uint32_t *counter; };
static int replay_stale(void *ctx, int h, uint32_t counter) {
    return ((struct replay_ctx *)ctx)->counter[h] > counter;
}
int ora_heap_replay(int n_ops, const int32_t *op, const int32_t *h, const double *q, const double *r, int n_handlers,
                    int first_push_ties, int32_t *out) {
    ora_heap H = {0};
    struct replay_ctx ctx;
    ctx.counter = (uint32_t *)calloc((size_t)n_handlers, sizeof(uint32_t));
    if (!ctx.counter) return -1;
    for (int i = 0; i < n_ops; i++) {
        out[i] = -2;
        if (op[i] == 0) oheap_insert(&H, q[i], r[i], h[i], ctx.counter[h[i]]);
        else if (op[i] == 1) ctx.counter[h[i]]++;
        else out[i] = oheap_root(&H, replay_stale, &ctx, first_push_ties);
    }
    free(H.a);
    free(ctx.counter);
    return 0;
}
