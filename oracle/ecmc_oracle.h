/*
 * ecmc_oracle.h -- CPU restatement of JeLLyFysh's event-chain Monte Carlo inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests, the smoke
 * check and the CPU-baseline leg of the benchmark load it to compare against the HIP
 * product path.  Nothing in the product package may include, link or call it.
 *
 * It restates, in plain C99 with IEEE double arithmetic and no FMA contraction, the
 * following reference files (paths relative to the reference's jellyfysh/ package):
 *   mediator/single_process_mediator.py:84-156       (leg loop)
 *   activator/tag_activator.py:166-316               (create / trash by tag, LIFO pools)
 *   activator/internal_state/single_active_cell_occupancy.py:95-237
 *   activator/internal_state/cell_occupancy/cells/cuboid_cells.py:54-297
 *   activator/internal_state/cell_occupancy/cells/cuboid_periodic_cells.py:74-207
 *   activator/tagger/excluded_cells_tagger.py:107-132, cell_boundary_tagger.py:147-149
 *   setting/hypercubic_setting.py:101-191            (Python float % semantics)
 *   base/time.py:87-182                              (Time = divmod quotient/remainder)
 *   base/vectors.py:29-214
 *   event_handler/two_leaf_unit_event_handler.py:105-154
 *   event_handler/two_leaf_unit_bounding_potential_event_handler.py:112-168
 *   event_handler/abstracts/event_handler_with_bounding_potential.py:75-101
 *   event_handler/cell_boundary_event_handler.py:98-173
 *   event_handler/abstracts/end_of_chain_event_handler.py:80-187
 *   event_handler/single_independent_active_periodic_direction_end_of_chain_event_handler.py:87-237
 *   event_handler/initial_chain_start_of_run_event_handler.py:97-131
 *   event_handler/final_time_end_of_run_event_handler.py:51-103
 *   event_handler/fixed_interval_sampling_event_handler.py:84-109
 *   potential/hard_sphere_potential.py:65-99, inverse_power_potential.py:71-179,
 *   potential/lennard_jones_potential.py:63-135, potential/abstracts.py:80-581,
 *   potential/displaced_even_power_potential.py:73-142,
 *   potential/inverse_power_coulomb_bounding_potential/...c:53-139,
 *   potential/merged_image_coulomb_potential/...c:77-274
 *   scheduler/list_scheduler.py:116-175, scheduler/heap_scheduler/heap.c:156-304
 * plus CPython's random module (MT19937 init_by_array, random(), expovariate, uniform,
 * randint via getrandbits) that the reference draws from.
 *
 * Math modes: ORA_MATH_LIBM calls glibc's pow/log/exp/erfc exactly where the reference
 * (CPython `**`, math.log, C libm) does; ORA_MATH_CR replaces pow/log with quad-precision
 * (libquadmath) evaluations rounded once to double, i.e. correctly rounded results.
 */
#ifndef ECMC_ORACLE_H
#define ECMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_MATH_LIBM = 0, ORA_MATH_CR = 1 };

/* ---- CPython-compatible Mersenne twister ------------------------------------------ */
typedef struct { uint32_t mt[624]; int32_t mti; uint64_t words; } ora_mt;
void ora_mt_seed(ora_mt *r, uint64_t seed);
uint32_t ora_mt_u32(ora_mt *r);
double ora_random(ora_mt *r);
double ora_expovariate(ora_mt *r, double lambd, int math_mode);
double ora_uniform(ora_mt *r, double a, double b);
int64_t ora_randint(ora_mt *r, int64_t a, int64_t b);

/* ---- Python float semantics --------------------------------------------------------- */
double ora_py_mod(double x, double y);
void ora_py_divmod(double x, double y, double *q, double *r);
void ora_time_add(double q, double r, double dt, double *oq, double *or_);
double ora_time_sub(double q1, double r1, double q2, double r2);

/* ---- math switch ---------------------------------------------------------------------- */
double ora_pow(double x, double y, int math_mode);
double ora_log(double x, int math_mode);
void ora_log_n(int64_t n, const double *x, double *out, int math_mode);
void ora_separation_norms(int dim, int64_t n, const double *pos, double L, double *out, int math_mode);
void ora_pow_n(int64_t n, const double *x, const double *y, double *out, int math_mode);

/* ---- potentials (scalar, reference argument conventions) --------------------------- */
/* HardSpherePotential.displacement(velocity, separation) */
double ora_hs_displacement(int dim, const double *velocity, const double *separation, double radius);
/* HardDipolePotential.displacement(velocity, separation) */
double ora_hd_displacement(int dim, const double *velocity, const double *separation, double minimum,
                           double maximum);
/* InversePowerPotential standard_velocity_* with direction of motion `dir` and speed. */
double ora_ip_displacement(int dim, int dir, double speed, const double *separation, double c1, double c2,
                           double power, double prefactor, double potential_change, int math_mode);
double ora_ip_derivative(int dim, int dir, double speed, const double *separation, double c1, double c2,
                         double power, double prefactor, int math_mode);
double ora_ip_potential(int dim, const double *separation, double charge_product, double power, double prefactor,
                        int math_mode);
/* Truncated and shifted inverse power (no reference counterpart; cutoff rc). */
double ora_ipt_displacement(int dim, int dir, double speed, const double *separation, double c1, double c2,
                            double power, double prefactor, double cutoff, double potential_change, int math_mode);
/* LennardJonesPotential (Mexican hat) */
double ora_lj_displacement(int dim, int dir, double speed, const double *separation, double prefactor,
                           double sigma, double potential_change, int math_mode);
double ora_lj_derivative(int dim, int dir, double speed, const double *separation, double prefactor,
                         double sigma, int math_mode);
/* DisplacedEvenPowerPotential (Mexican hat) */
double ora_dep_displacement(int dim, int dir, double speed, const double *separation, double r0, int power,
                            double prefactor, double potential_change, int math_mode);
double ora_dep_derivative(int dim, int dir, double speed, const double *separation, double r0, int power,
                          double prefactor, int math_mode);
/* InversePowerCoulombBoundingPotential (x is the direction of motion; use permuted separations) */
double ora_ipcb_derivative(double prefactor_product, double sx, double sy, double sz);
double ora_ipcb_displacement(double prefactor_product, double sx, double sy, double sz, double potential_change,
                             double system_length);
/* MergedImageCoulombPotential derivative along +x (the C kernel; no prefactor/charges) */
typedef struct ora_ewald ora_ewald;
ora_ewald *ora_ewald_create(int fourier_cutoff, int position_cutoff, double alpha, double system_length);
void ora_ewald_destroy(ora_ewald *e);
double ora_ewald_derivative(const ora_ewald *e, double sx, double sy, double sz);

/* ---- cells ---------------------------------------------------------------------------- */
typedef struct ora_cells ora_cells;
ora_cells *ora_cells_create(int dim, double system_length, const int *cells_per_side, int neighbor_layers);
void ora_cells_destroy(ora_cells *c);
int ora_cells_count(const ora_cells *c);
void ora_cells_bounds(const ora_cells *c, int cell, double *cell_min, double *cell_max);
int ora_cells_position_to_cell(const ora_cells *c, const double *position);
int ora_cells_nearby(const ora_cells *c, int cell, int *out); /* canonical order, returns count */
int ora_cells_neighbor(const ora_cells *c, int cell, int direction, int positive);
/* cuboid_periodic_cells.py:182-207: the cell holding the centre of `cell` shifted by the
 * lower corner of `rel` */
int ora_cells_translate(const ora_cells *c, int cell, int rel, double system_length);

/* ---- cell bounding potential (cell_bounding_potential.py:70-126 with
 * inner_point_estimator.py:108-171) ---------------------------------------------------------
 * For every cell c not among the nearby cells of the zero cell and every direction d:
 * out[(c * 3 + d) * 2] = upper and out[(c * 3 + d) * 2 + 1] = lower bound of the derivative
 * prefactor * 1.0 * target_charge * ewald_derivative over the (pps + 1)^3 grid spanning
 * [c.min - zero.max, c.max - zero.min] (separations corrected for periodic boundaries),
 * scaled by the estimator prefactor and clipped by the empirical bound; NaN for nearby
 * cells.  3-D only. */
void ora_cell_bound_table(const ora_cells *cells, const ora_ewald *ewald, double potential_prefactor,
                          double target_charge, int points_per_side, double estimator_prefactor,
                          double empirical_bound, double system_length, double *out);

/* ---- Walker's alias table (walker.py:53-130) ------------------------------------------
 * Built from m rates in the given order.  Entry i holds (item0[i], rate0[i]) and, when
 * item1[i] >= 0, the alias item1[i] (its rate is mean - rate0[i]); entries without an alias
 * carry rate0 = mean.  ora_walker_build returns the number of entries (m) or -1 when a rate is
 * negative or the leftovers fail the reference's 1e-6 check. */
typedef struct {
    int m;
    double total, mean;
    int32_t *item0, *item1;
    double *rate0;
} ora_walker;
int ora_walker_build(ora_walker *w, int m, const int32_t *items, const double *rates);
void ora_walker_free(ora_walker *w);
/* sample_cell: random.choice over the entries, then uniform(0, mean) <= rate0 keeps item0 */
int32_t ora_walker_sample(const ora_walker *w, ora_mt *r);

/* ---- the chain engine ---------------------------------------------------------------- */
enum {
    ORA_TAG_PAIR_CELLS = 0,   /* ExcludedCellsTagger -> pair event handler */
    ORA_TAG_PAIR_ALL = 1,     /* every other unit in index order -> pair event handler */
    ORA_TAG_CELL_BOUNDARY = 2,
    ORA_TAG_SAMPLING = 3,
    ORA_TAG_END_OF_CHAIN = 4,
    ORA_TAG_END_OF_RUN = 5,
    ORA_TAG_START_OF_RUN = 6,
    ORA_TAG_SURPLUS = 7,      /* SurplusCellsTagger: the surplus units in dict order -> pair handler */
    ORA_TAG_CELL_BOUNDING = 8, /* CellBoundingPotentialTagger: occupants of the non-nearby cells */
    ORA_TAG_CELL_VETO = 9,     /* CellVetoTagger: one handler per leg for the active unit */
    ORA_TAG_FACTOR_MAP = 10    /* FactorTypeMapInStateTagger over two-leaf factors (see ora_tagger.fmap_*) */
};
/* ORA_EH_CELL_BOUNDING: TwoLeafUnitCellBoundingPotentialEventHandler with a merged-image
 * Coulomb potential (par) and a CellBoundingPotential from an InnerPointEstimator
 * (bpar: estimator prefactor, points_per_side, target_charge, empirical_bound). */
/* ORA_EH_CELL_VETO: LeafUnitCellVetoEventHandler (leaf_unit_cell_veto_event_handler.py:31-149,
 * cell_veto_event_handler.py:93-238) with the merged-image Coulomb potential (par) and an
 * InnerPointEstimator (bpar as for ORA_EH_CELL_BOUNDING). */
enum { ORA_EH_TWO_LEAF = 0, ORA_EH_TWO_LEAF_BOUNDING = 1, ORA_EH_CELL_BOUNDING = 2, ORA_EH_CELL_VETO = 3 };
enum {
    ORA_POT_HARD_SPHERE = 0,      /* par: radius */
    ORA_POT_INVERSE_POWER = 1,    /* par: power, prefactor */
    ORA_POT_LENNARD_JONES = 2,    /* par: prefactor, sigma */
    ORA_POT_INVERSE_POWER_TRUNC = 3, /* par: power, prefactor, cutoff */
    ORA_POT_MERGED_IMAGE_COULOMB = 4, /* par: alpha, fourier_cutoff, position_cutoff, prefactor */
    ORA_POT_IPCB = 5,             /* par: prefactor */
    ORA_POT_DISPLACED_EVEN_POWER = 6, /* par: equilibrium_separation, power, prefactor */
    ORA_POT_HARD_DIPOLE = 8       /* par: minimum_separation, maximum_separation */
};
enum { ORA_SCHED_LIST = 0, ORA_SCHED_HEAP = 1 };
/* Events of equal time: ORA_TIES_SCHEDULER ranks them as the configured scheduler does
 * (heap.c's structural order, or the list scheduler's first push); ORA_TIES_FIRST_PUSH ranks
 * them by push order under either scheduler (the device chain's rule). */
enum { ORA_TIES_SCHEDULER = 0, ORA_TIES_FIRST_PUSH = 1 };

#define ORA_MAX_TAGGERS 16
#define ORA_MAX_LEAVES 4
#define ORA_MAX_FACTOR_MAP 8
typedef struct {
    int32_t kind;
    int32_t event_handler;
    int32_t potential, bounding_potential;
    double par[4], bpar[4];
    int32_t use_charge;
    int32_t n_create, create[ORA_MAX_TAGGERS];
    int32_t n_trash, trash[ORA_MAX_TAGGERS];
    int32_t n_handlers;          /* pool size (number_event_handlers) */
    double value;                /* sampling_interval | chain_time | end_of_run_time | speed */
    int32_t ivalue;              /* first_event_time_zero | initial_direction_of_motion */
    int32_t ivalue2;             /* initial_active_identifier (leaf units: root * leaves_per_root + leaf) */
    /* ORA_TAG_FACTOR_MAP (factor_type_maps.py:263-356, factor_type_map_in_state_tagger.py:83-107):
     * for leaf number k of the active unit, fmap_n[k] two-leaf factors whose other leaf has
     * number fmap_target[k][e] in the numbering of two composite objects; local factors lie
     * within the active unit's own object, non-local ones are yielded for every other object in
     * identifier order.  The tagger yields them through a set: the order here is the insertion
     * order (the order of an insertion-ordered set, as PyPy's). */
    int32_t fmap_local;
    int32_t fmap_n[ORA_MAX_LEAVES];
    int32_t fmap_target[ORA_MAX_LEAVES][ORA_MAX_FACTOR_MAP];
} ora_tagger;

typedef struct {
    int32_t dim, n;
    double system_length, beta;
    int32_t cells_per_side[3];
    int32_t neighbor_layers;
    int32_t max_occupants;       /* <= 0: unbounded */
    int32_t scheduler;
    int32_t math_mode;
    int32_t rng_skip;            /* words drawn before the run (the random input handler's draws) */
    int32_t ties;                /* ORA_TIES_* */
    int32_t leaves_per_root;     /* leaf units per composite point object (0 or 1: point masses) */
    uint64_t seed;
    int32_t n_taggers;
    ora_tagger taggers[ORA_MAX_TAGGERS];
} ora_config;

/* One committed leg.  `tagger` is the tagger index of the event handler that committed. */
typedef struct {
    int32_t tagger;
    int32_t flags;      /* bit0: lifting happened (velocity changed owner) */
    int32_t active;     /* active unit before the event */
    int32_t target;     /* pair target / new active chosen by end-of-chain / -1 */
    double tq, tr;      /* event time (Time quotient, remainder) */
    double pos[3];      /* committed position of the previously active unit */
    int32_t ncand;      /* pair candidates created at the start of this leg */
    uint32_t chash;     /* sum over them of an order-dependent hash of (rank, target) */
} ora_event;

typedef struct ora_engine ora_engine;
ora_engine *ora_engine_create(const ora_config *cfg, const double *positions, const double *charges);
void ora_engine_destroy(ora_engine *e);
/* Run until `max_events` legs were committed or the end-of-run event fired.  Returns the
 * number of legs written to `log` (may be NULL), or a negative error code. */
int64_t ora_engine_run(ora_engine *e, int64_t max_events, ora_event *log);
int ora_engine_done(const ora_engine *e);
int ora_engine_error(const ora_engine *e, char *buf, int buflen);
void ora_engine_positions(const ora_engine *e, double *out);
int32_t ora_engine_active(const ora_engine *e, int32_t *direction, double *tq, double *tr);
/* Estimator sums (pair-lift separations along the direction of motion, chain displacement,
 * counts of committed pair/cell-boundary/end-of-chain legs, lifts). */
void ora_engine_estimators(const ora_engine *e, double *out8);
/* Candidate trace of the last leg that created pair candidates: count and the
 * target ids in push order (up to cap). */
int32_t ora_engine_last_candidates(const ora_engine *e, int32_t *out, int32_t cap);
/* Number of random words consumed so far. */
int64_t ora_engine_rng_words(const ora_engine *e);

/* heap.c on its own: a sequence of pushes, trashes and root requests (see the engine source) */
int ora_heap_replay(int n_ops, const int32_t *op, const int32_t *h, const double *q, const double *r, int n_handlers,
                    int first_push_ties, int32_t *out);

/* ---- batched replicas (the product's batched form; see ecmc_oracle_batched.c) ---------- */
void ora_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
typedef struct {
    int32_t n, n_factors;
    double system_length, beta, cutoff, chain_length;
    double power[2], prefactor[2];
    uint64_t seed;
    uint32_t replica;
    int32_t math_mode;   /* ORA_MATH_CR: what the device reproduces; ORA_MATH_LIBM: faster */
} ora_batched_config;
typedef struct {
    int32_t active, direction;
    double chain_left;
    uint64_t leg;
    double est[8];   /* lift separation sum, displacement sum, pair, boundary, lifts, chain ends */
} ora_batched_state;
typedef struct {
    int32_t kind;     /* 0 pair, 1 cell boundary, 2 end of chain */
    int32_t active, target, factor;
    double displacement, x;   /* event displacement, committed coordinate along the direction */
} ora_batched_event;
int64_t ora_batched_run(const ora_batched_config *cfg, double *positions, ora_batched_state *st, int64_t legs,
                        ora_batched_event *log);

/* Plain reference-wiring helpers for the benchmark configurations. */
void ora_config_hard_disks(ora_config *c, int n, double system_length, double radius, int cells_per_side,
                           double chain_time, double end_time, uint64_t seed);
void ora_config_inverse_power_pairs(ora_config *c, int dim, int n, double system_length, double beta,
                                    int cells_per_side, int n_factors, const double *powers,
                                    const double *prefactors, double cutoff, double chain_time, double end_time,
                                    uint64_t seed, int math_mode);

/* Independent replicas on host threads (the CPU baseline).  Each thread copies `cfg` with
 * seed+i and runs `events_per_replica` legs.  Returns total committed legs; wall time in *seconds. */
int64_t ora_run_threads(const ora_config *cfg, const double *positions, const double *charges, int n_threads,
                        int64_t events_per_replica, double *seconds);

#ifdef __cplusplus
}
#endif
#endif
