/*
 * jellyfysh_amd.h -- C ABI of the MI355X ECMC hot path (libjellyfysh_amd.so).
 *
 * Plain C: every argument is a pointer, a size or a plain struct; no torch or HIP types.
 * Functions return 0 on success and a negative ecmc_status otherwise; ecmc_last_error()
 * gives the message of the last failure on the calling thread.  "No event" is +INF, as in
 * the reference (a time that the heap scheduler never stores, heap_scheduler.py:139).
 * Device pointers (`d_` prefix) are HIP device memory; a `stream` argument is a
 * hipStream_t passed as void* (NULL = the default stream).
 *
 * Reference interfaces replaced (paths relative to the reference's jellyfysh/ package):
 *   ecmc_system_*          the mediator leg loop with its activator, cell occupancy,
 *                          taggers, event handlers and scheduler for one chain
 *                          (mediator/single_process_mediator.py:84-156,
 *                          activator/tag_activator.py:166-316,
 *                          activator/internal_state/single_active_cell_occupancy.py:95-237,
 *                          scheduler/heap_scheduler/heap.c:156-304 + list_scheduler.py:116-175)
 *   ecmc_pair_displacement InvertiblePotential.displacement(velocity, separation, charge_one,
 *                          charge_two, potential_change) (potential/potential.py:216-253) for
 *                          the hard-sphere, hard-dipole, inverse-power, Lennard-Jones,
 *                          displaced-even-power and Coulomb-bounding potentials, batched
 *   ecmc_pair_derivative   Potential.derivative(velocity, separation, charges)
 *                          (potential/potential.py:153-181), batched; for the merged-image
 *                          Coulomb potential this replaces merged_image_coulomb_potential.c:205-274
 *   ecmc_math_log/pow      the libm log / pow the reference's Python calls through `**` and
 *                          math.log (correctly rounded here)
 *   ecmc_mt_*              CPython random (MT19937, random(), expovariate, randint) as the
 *                          reference draws from it (two_leaf_unit_event_handler.py:98-103,
 *                          single_independent_active_periodic_direction_end_of_chain_event_handler.py:229)
 */
#ifndef JELLYFYSH_AMD_H
#define JELLYFYSH_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ECMC_OK = 0,
    ECMC_ERR_ARGUMENT = -1,
    ECMC_ERR_HIP = -2,
    ECMC_ERR_CONFIGURATION = -3, /* base.exceptions.ConfigurationError */
    ECMC_ERR_SCHEDULER = -4,     /* base.exceptions.SchedulerError */
    ECMC_ERR_TAG_ACTIVATOR = -5, /* base.exceptions.TagActivatorError */
    ECMC_ERR_STATE = -6,         /* a position left the cell system or an occupancy list overflowed */
    ECMC_ERR_NUMERIC = -7        /* the Python code would raise ValueError (sqrt of a negative number) */
} ecmc_status;

/* Tagger kinds (the modules of activator/tagger/) */
enum {
    ECMC_TAG_PAIR_CELLS = 0,     /* ExcludedCellsTagger: active x occupants of the nearby cells */
    ECMC_TAG_PAIR_ALL = 1,       /* every other unit in identifier order */
    ECMC_TAG_CELL_BOUNDARY = 2,  /* CellBoundaryTagger + CellBoundaryEventHandler */
    ECMC_TAG_SAMPLING = 3,       /* NoInStateTagger + FixedIntervalSamplingEventHandler */
    ECMC_TAG_END_OF_CHAIN = 4,   /* ActiveGlobalStateInStateTagger + periodic-direction end of chain */
    ECMC_TAG_END_OF_RUN = 5,     /* NoInStateTagger + FinalTimeEndOfRunEventHandler */
    ECMC_TAG_START_OF_RUN = 6,   /* NoInStateTagger + InitialChainStartOfRunEventHandler */
    ECMC_TAG_SURPLUS = 7,        /* SurplusCellsTagger: the surplus units of a bounded occupancy */
    ECMC_TAG_CELL_BOUNDING = 8,  /* CellBoundingPotentialTagger: occupants of the non-nearby cells */
    ECMC_TAG_CELL_VETO = 9,      /* CellVetoTagger (cell_veto_tagger.py:31-155): the active unit once */
    ECMC_TAG_FACTOR_MAP = 10     /* FactorTypeMapInStateTagger (factor_type_map_in_state_tagger.py:83-107) over
                                    two-leaf factors: the leaf units that the active leaf unit meets through the
                                    tagger's factor type map (activator/tagger/factor_type_maps.py:325-356), other
                                    composite objects in identifier order, map entries in file order */
};
/* Event handlers of pair taggers */
/* ECMC_EH_CELL_BOUNDING: TwoLeafUnitCellBoundingPotentialEventHandler
 * (two_leaf_unit_cell_bounding_potential_event_handler.py:113-211) with the merged-image
 * Coulomb potential (par) and ECMC_POT_CELL_BOUNDING (bpar). */
/* ECMC_EH_CELL_VETO: LeafUnitCellVetoEventHandler (leaf_unit_cell_veto_event_handler.py:31-149,
 * cell_veto_event_handler.py:93-238): Walker tables over the same InnerPointEstimator bounds
 * (bpar as for ECMC_EH_CELL_BOUNDING), one sampled target cell per leg, thinning against the
 * merged-image Coulomb potential (par). */
enum { ECMC_EH_TWO_LEAF = 0, ECMC_EH_TWO_LEAF_BOUNDING = 1, ECMC_EH_CELL_BOUNDING = 2, ECMC_EH_CELL_VETO = 3 };
/* Potentials (the modules of potential/) */
enum {
    ECMC_POT_HARD_SPHERE = 0,              /* par: radius */
    ECMC_POT_INVERSE_POWER = 1,            /* par: power, prefactor */
    ECMC_POT_LENNARD_JONES = 2,            /* par: prefactor, characteristic_length */
    ECMC_POT_INVERSE_POWER_TRUNC = 3,      /* par: power, prefactor, cutoff (no reference counterpart) */
    ECMC_POT_MERGED_IMAGE_COULOMB = 4,     /* par: alpha, fourier_cutoff, position_cutoff, prefactor */
    ECMC_POT_INVERSE_POWER_COULOMB_BOUND = 5, /* par: prefactor */
    ECMC_POT_DISPLACED_EVEN_POWER = 6,     /* par: equilibrium_separation, power, prefactor */
    ECMC_POT_CELL_BOUNDING = 7,            /* CellBoundingPotential + InnerPointEstimator; par: estimator
                                              prefactor, points_per_side, target_charge, empirical_bound */
    ECMC_POT_HARD_DIPOLE = 8               /* HardDipolePotential; par: minimum_separation, maximum_separation */
};
enum { ECMC_SCHED_LIST = 0, ECMC_SCHED_HEAP = 1 };

#define ECMC_MAX_TAGGERS 16
#define ECMC_MAX_LEAVES 4        /* leaf units per composite point object */
#define ECMC_MAX_FACTOR_MAP 8    /* factor type map entries per leaf number */

/* One tagger of the TagActivator with its event handler (an INI section pair). */
typedef struct {
    int32_t kind;
    int32_t event_handler;
    int32_t potential, bounding_potential;
    double par[4], bpar[4];
    int32_t use_charge;
    int32_t n_create, create[ECMC_MAX_TAGGERS]; /* tagger indices, in INI order */
    int32_t n_trash, trash[ECMC_MAX_TAGGERS];
    int32_t n_handlers;                         /* number_event_handlers */
    double value;   /* sampling_interval | chain_time | end_of_run_time | speed */
    int32_t ivalue; /* first_event_time_zero | initial_direction_of_motion */
    int32_t ivalue2;/* initial_active_identifier (leaf units: root * leaves_per_root + leaf) */
    /* ECMC_TAG_FACTOR_MAP: the factor type map of the tagger's factor for two composite objects
     * of leaves_per_root leaves each (factor_type_maps.py:37-88): for the active unit's leaf
     * number k, fmap_n[k] two-leaf factors whose other leaf has the number fmap_target[k][e] in
     * the two-object numbering (0 .. leaves_per_root - 1 this object, the rest the other one);
     * fmap_local: the factor lies within one composite object */
    int32_t fmap_local;
    int32_t fmap_n[ECMC_MAX_LEAVES];
    int32_t fmap_target[ECMC_MAX_LEAVES][ECMC_MAX_FACTOR_MAP];
} ecmc_tagger_desc;

/* A run configuration: HypercubicSetting + CuboidPeriodicCells + SingleActiveCellOccupancy
 * + scheduler + taggers, and the seed of the random stream (random.seed(seed)).  With composite
 * point objects (leaves_per_root > 1, TreeStateHandler with two node levels) the units are the
 * leaf units, which alone carry velocities here; the new active unit of an end of chain is then
 * randint(root) followed by randint(leaf)
 * (single_independent_active_periodic_direction_end_of_chain_event_handler.py:176-237).  rng_skip
 * words of that stream are drawn before the run starts, as a random input handler draws the
 * initial positions from it (random_input_handler.py:79-92: two words per coordinate). */
typedef struct {
    int32_t dim, n;
    double system_length, beta;
    int32_t cells_per_side[3];
    int32_t neighbor_layers;
    int32_t max_occupants;   /* <= 0: unbounded */
    int32_t scheduler;
    int32_t rng_skip;
    int32_t leaves_per_root; /* leaf units per composite point object (0 or 1: point masses); unit
                                root * leaves_per_root + leaf, n a multiple of it */
    uint64_t seed;
    int32_t n_taggers;
    ecmc_tagger_desc taggers[ECMC_MAX_TAGGERS];
} ecmc_system_desc;

/* One committed leg.  `ncand`/`chash` describe the pair candidates created right after
 * this event (count and an order-dependent hash of the target identifiers). */
typedef struct {
    int32_t tagger;
    int32_t flags;      /* bit0: the velocity changed owner (lifting) */
    int32_t active;     /* active unit before the event */
    int32_t target;     /* pair target, new active of an end of chain, or -1 */
    double tq, tr;      /* event time as Time(quotient, remainder) */
    double pos[3];      /* committed position of the previously active unit */
    int32_t ncand;
    uint32_t chash;
} ecmc_event;

typedef struct ecmc_system ecmc_system;

const char *ecmc_last_error(void);
/* Number of HIP devices visible (0 when none). */
int ecmc_device_count(void);

/* ---- one chain on the GPU (plain form) ---------------------------------------------------
 * Create uploads the positions (n x dim, row-major) and charges (n, may be NULL) and
 * prepares cells and occupancy exactly as the reference's initialisation does. */
int ecmc_system_create(const ecmc_system_desc *desc, const double *positions, const double *charges,
                       ecmc_system **out);
void ecmc_system_destroy(ecmc_system *sys);
/* Commit up to max_events legs (stops after the end-of-run event).  Events are copied to
 * host_log when it is not NULL (room for max_events records). */
int ecmc_system_run(ecmc_system *sys, int64_t max_events, ecmc_event *host_log, int64_t *n_committed);
/* Bit t set: ecmc_system_run returns right after committing an event of tagger t (e.g. a
 * sampling event, so that the caller can read the state at the sampling time). */
int ecmc_system_set_stop_mask(ecmc_system *sys, int32_t mask);
int ecmc_system_positions(const ecmc_system *sys, double *positions_out);
int ecmc_system_active(const ecmc_system *sys, int32_t *active, int32_t *direction, double *tq, double *tr,
                       int32_t *done);
/* est[0] = sum of lifting separations along the direction of motion, est[1] = chain
 * displacement of completed chains, est[2] = pair legs, est[3] = cell-boundary legs,
 * est[4] = lifts, est[5] = end-of-chain legs, est[6] = rejected thinning candidates. */
int ecmc_system_estimators(const ecmc_system *sys, double *est8);
/* Number of MT words drawn so far. */
int ecmc_system_rng_words(const ecmc_system *sys, int64_t *words);
/* Kernel time of the last ecmc_system_run in milliseconds (HIP events on its stream). */
int ecmc_system_last_kernel_ms(const ecmc_system *sys, double *ms);
/* The dynamic LDS of the system's leg-loop launches: bytes, and whether its positions and its
 * random-word ring live there for a launch (1) or stay in global memory (0).  No reference
 * counterpart: the device's memory placement, for diagnostics. */
int ecmc_system_launch_lds(const ecmc_system *sys, int64_t *bytes, int32_t *positions, int32_t *ring);
/* The leg-loop kernel that runs the system alone: *features gets the features its taggers use
 * (bit p for the two-leaf pair potential ECMC_POT_p inverted exactly, bit 9 the merged-image
 * Coulomb factors through the inverse-power Coulomb bound, bit 10 factor-type maps, bit 11 cell
 * bounding, bit 12 the cell veto), *variant the features of the specialized kernel launched for them (its
 * template argument in the kernel's name, ecmc::chain_kernel<dim, variant>).  A launch of
 * several systems runs the first variant covering all of them.  No reference counterpart: which
 * compiled code runs, for diagnostics and profiles. */
int ecmc_system_kernel_variant(const ecmc_system *sys, int32_t *features, int32_t *variant);
/* The same for a descriptor, before any system exists (no device call).  No reference
 * counterpart. */
int ecmc_desc_kernel_variant(const ecmc_system_desc *desc, int32_t *features, int32_t *variant);
/* Independent systems advanced together: one launch, one workgroup per system, each system
 * committing up to max_events legs exactly as ecmc_system_run would (no event logs).  All
 * systems have one dimension and appear once; n_committed (count entries, may be NULL) gets
 * each system's committed legs; the launch runs on the first system's stream, whose kernel
 * time every system then reports.  Replaces, for a set of independent runs, one
 * SingleProcessMediator.run per system (mediator/single_process_mediator.py:84-156). */
int ecmc_systems_run(ecmc_system *const *systems, int32_t count, int64_t max_events, int64_t *n_committed);
/* ecmc_systems_run without any allocation (no hipMalloc, hipFree or other allocating call): the
 * launch arguments go into a fixed device array of the library, sized for
 * ECMC_SYSTEMS_NOALLOC_MAX systems (512 B each; every GPU the library runs on holds its own
 * copy until the process ends), staged through a fixed host array.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  count above the maximum is rejected with ECMC_ERR_ARGUMENT; the
 * systems must live on the current device.  Calls that overlap on one GPU share the array and
 * would race: the library serialises the calls of one process with a lock. */
#define ECMC_SYSTEMS_NOALLOC_MAX 1024
int ecmc_systems_run_noalloc(ecmc_system *const *systems, int32_t count, int64_t max_events, int64_t *n_committed);

/* ---- the leg step by step: a set-level event handler per tagger -------------------------------
 * The host keeps the mediator loop, the scheduler and the tag activator's create/trash lists
 * (mediator/single_process_mediator.py:84-156, mediator/mediator.py:142-211); the state and the
 * candidate sets stay on the device.  Each tagger acts as one event handler for its whole set
 * of factors, the in-tree pattern of handlers that return (time, arguments)
 * (event_handler/event_handler.py:80-125; event_handler/abstracts/cell_veto_event_handler.py:162-238
 * with mediator/mediator.py:265-292):
 *   ecmc_system_begin_leg        the internal-state update before a leg's creates (the cell
 *                                occupancy follows the active unit, tag_activator.py:205-226);
 *   ecmc_system_send_event_time  send_event_time of tagger `tagger`: every candidate of its set
 *                                is evaluated on the device (random draws in the reference's
 *                                order); returns the earliest time as Time(tq, tr) and its target
 *                                unit (the new active unit for an end of chain, -1 for events
 *                                without one); valid = 0 (time +INF) when the set is empty;
 *   ecmc_system_send_out_state   send_out_state of the stored event of `tagger` and its
 *                                insertion into the state (thinning included); the committed
 *                                event record goes to *event.
 * The host's scheduler decides which tagger's event comes next; trashing is the host's alone. */
int ecmc_system_begin_leg(ecmc_system *sys);
int ecmc_system_send_event_time(ecmc_system *sys, int32_t tagger, double *tq, double *tr, int32_t *target,
                                int32_t *valid);
int ecmc_system_send_out_state(ecmc_system *sys, int32_t tagger, ecmc_event *event);

/* ---- batched independent replicas (no reference counterpart as a whole) ------------------
 * Each replica is a reference chain (straight chains of chain_length cycling +x,+y,+z,
 * exact pair-factor inversion, cell-boundary events of CuboidPeriodicCells, a random new
 * active unit at each chain end) of LJ-type truncated-and-shifted inverse-power factors
 * (powers 6 and/or 12) in a cubic box, with Philox4x32-10 draws keyed by (replica, seed)
 * and counter (leg, unit).  Candidate displacements are bounded in FP32 first; only the
 * few whose bound can decide the leg are inverted exactly in FP64. */
typedef struct {
    int32_t n_replicas;
    int32_t n_factors;
    double beta, cutoff, chain_length;
    double power[2], prefactor[2];
    uint64_t seed;
} ecmc_batch_desc;

typedef struct {
    int32_t kind;        /* 0 pair (lift), 1 cell boundary, 2 end of chain */
    int32_t active, target, factor;
    double displacement; /* event displacement along the chain */
    double x;            /* committed coordinate of the moved unit along the direction */
} ecmc_batch_event;

typedef struct ecmc_batch ecmc_batch;

/* positions: replicas concatenated, n_i x 3 row-major each, inside [0, L_i). */
int ecmc_batch_create(const ecmc_batch_desc *desc, const int32_t *n_per_replica, const double *system_length,
                      const double *positions, ecmc_batch **out);
void ecmc_batch_destroy(ecmc_batch *batch);
/* Size classes: replicas with n <= class_max_n[c] (first such c) run with class_wpr[c]
 * wavefronts per replica; one persistent kernel per class, all launched together. */
int ecmc_batch_set_classes(ecmc_batch *batch, int32_t n_classes, const int32_t *class_max_n, const int32_t *class_wpr);
/* 0: FP32 filter (default); 1: every candidate inverted in FP64; 2: filter plus audit of
 * every bound against the exact value (violations counted, see ecmc_batch_audit). */
int ecmc_batch_set_mode(ecmc_batch *batch, int32_t mode);
/* Record the first `legs_per_replica` legs of every replica (0 = off). */
int ecmc_batch_set_log(ecmc_batch *batch, int32_t legs_per_replica);
/* One step: every replica advances legs_per_replica legs (synchronous). */
int ecmc_batch_run(ecmc_batch *batch, int32_t legs_per_replica);
/* The step captured once into a hipGraph and replayed n_steps times (synchronous). */
int ecmc_batch_run_graph(ecmc_batch *batch, int32_t legs_per_replica, int32_t n_steps);
int ecmc_batch_last_ms(const ecmc_batch *batch, double *ms);
/* Limit every size class's persistent grid to max_blocks workgroups (0 = as many as fit), e.g.
 * to leave room for other work on the device.  A replica's legs run in chunks, and any block
 * may take the next chunk, so results do not depend on the limit.  No reference counterpart. */
int ecmc_batch_set_max_blocks(ecmc_batch *batch, int32_t max_blocks);
/* Class timing: when on, ecmc_batch_run times every size class's kernel with events on the
 * class's own stream, and ecmc_batch_class_ms returns its kernel milliseconds and launches (1,
 * or 0 for an empty class) of the last run call; graph replays are not class-timed (launches
 * 0).  No reference counterpart: the reference has no batched form. */
int ecmc_batch_set_class_timing(ecmc_batch *batch, int32_t on);
int ecmc_batch_class_ms(const ecmc_batch *batch, int32_t cls, double *ms, int32_t *launches);
/* Per replica: active unit, direction, chain displacement left, legs done, est[8] (lift
 * separation sum, displacement sum, pair, boundary, lifts, chain ends) and stats[4]
 * (candidates, candidates past the geometric test, contenders, exact inversions). */
int ecmc_batch_states(const ecmc_batch *batch, int32_t *active, int32_t *direction, double *chain_left,
                      uint64_t *legs, double *est8, int64_t *stats4);
int ecmc_batch_positions(const ecmc_batch *batch, double *positions);
int ecmc_batch_log(const ecmc_batch *batch, ecmc_batch_event *host_log);
int ecmc_batch_audit(const ecmc_batch *batch, uint64_t *violations);
/* Snapshot and restore of a batch's state between steps: every replica's state (active unit,
 * direction, chain left, leg counter, which is the Philox counter, estimators, statistics) and
 * all positions, into a host buffer of *bytes bytes (buffer NULL: *bytes gets the size).  A
 * restore into a batch of the same description and replica sizes continues every replica as
 * the snapshotted batch would have.  Stands for the reference's dumping output handler and
 * resume.py (input_output_handler/output_handler/dumping_output_handler.py:70-90,
 * resume.py:63-109), for the batched form. */
int ecmc_batch_snapshot(const ecmc_batch *batch, void *buffer, int64_t *bytes);
int ecmc_batch_restore(ecmc_batch *batch, const void *buffer, int64_t bytes);

/* ---- batched per-pair evaluations on device memory ---------------------------------------
 * Row i uses direction dir[i] (velocity speed[i] along +dir), separation sep[3*i..3*i+dim)
 * (target minus active, minimum image), charges c1[i], c2[i] (NULL = 1.0) and potential
 * change dU[i] (ignored where the potential takes none). */
int ecmc_pair_displacement(int32_t potential, const double *par, int32_t dim, int64_t n, const int32_t *d_dir,
                           const double *d_speed, const double *d_sep, const double *d_c1, const double *d_c2,
                           const double *d_dU, double system_length, double *d_out, void *stream);
int ecmc_pair_derivative(int32_t potential, const double *par, int32_t dim, int64_t n, const int32_t *d_dir,
                         const double *d_speed, const double *d_sep, const double *d_c1, const double *d_c2,
                         double system_length, double *d_out, void *stream);
/* FP32 brackets lb <= t <= ub of the time displacement t that ecmc_pair_displacement returns for
 * the inverse-power potentials (ECMC_POT_INVERSE_POWER, _TRUNC; integer power 1..32) and for the
 * inverse-power Coulomb bound (ECMC_POT_INVERSE_POWER_COULOMB_BOUND, dim 3, par = {prefactor,
 * system_length}; inverse_power_coulomb_bounding_potential.c:85-139), with the
 * potential change dU = -log(1 - u) / beta of the random pair whose leading MT word is word[i]
 * (u = random() of the words (word[i], next word), potential/inverse_power_potential.py:96-179
 * with the expovariate draw of event_handler/two_leaf_unit_event_handler.py).  lb = +inf only
 * where t is +inf; ub = +inf where an event is not certain.  The chain kernel inverts exactly
 * only the candidates whose lb does not exceed the smallest ub of their set.  No reference
 * counterpart as an entry point: the brackets behind the candidate filter, exposed for tests. */
int ecmc_pair_time_bounds(int32_t potential, const double *par, int32_t dim, int64_t n, const int32_t *d_dir,
                          const double *d_speed, const double *d_sep, const double *d_c1, const double *d_c2,
                          const uint32_t *d_word, double beta, float *d_lb, float *d_ub, void *stream);

/* ---- merged-image Coulomb handle ------------------------------------------------------------
 * The opaque handle of merged_image_coulomb_potential.h:37-44 (construct_, copy_,
 * destroy_merged_image_coulomb_potential, estimated_size), its Fourier table resident on the
 * device that is current at creation.  ecmc_ewald_derivative is the C file's derivative
 * (merged_image_coulomb_potential.c:205-274) for n separations (n x 3, device memory, x along
 * the direction of motion as vectors.permutation_3d orders them), without prefactor and
 * charges, into d_out.  NULL / a negative status stands for the reference's MemoryError. */
typedef struct ecmc_ewald ecmc_ewald;
int ecmc_ewald_create(int32_t fourier_cutoff, int32_t position_cutoff, double alpha, double system_length,
                      ecmc_ewald **out);
int ecmc_ewald_copy(const ecmc_ewald *ewald, ecmc_ewald **out);
void ecmc_ewald_destroy(ecmc_ewald *ewald);
int ecmc_ewald_size(const ecmc_ewald *ewald, int64_t *bytes);
int ecmc_ewald_derivative(const ecmc_ewald *ewald, int64_t n, const double *d_sep, double *d_out, void *stream);

/* Plain device-to-device copy of `bytes` (a multiple of 16).  No reference counterpart: a
 * kernel of known traffic that scales hardware byte counters in profiles. */
int ecmc_copy_bytes(const void *d_src, void *d_dst, int64_t bytes, void *stream);

/* ---- pair energy and its backward kernel (no reference counterpart as a whole) -----------
 * U = 1/2 sum_{i != j in scope} U(s_ij) over minimum-image separations, for the inverse-power
 * (par: power, prefactor), truncated inverse-power (power, prefactor, cutoff <= L/2),
 * Lennard-Jones (prefactor, characteristic_length), displaced-even-power (equilibrium
 * separation, power, prefactor) and merged-image Coulomb (alpha, fourier_cutoff,
 * position_cutoff, prefactor) potentials.  The backward kernel gives dU/dx (n x 3: for unit
 * i along axis d the sum over j of the reference's derivative(e_d, s_ij, c_i, c_j),
 * potential/potential.py:153-181), dU/dpar (4, in the order of par; the Coulomb alpha and
 * cutoffs get 0) and, for charged terms, dU/dc (n).  Positions are n x 3 row-major in
 * [0, L); d_charge and d_molecule may be NULL when the term needs neither.  d_work is
 * device scratch of ecmc_energy_workspace_bytes() bytes; nothing is allocated. */
enum { ECMC_SCOPE_ALL = 0, ECMC_SCOPE_INTER = 1 /* different molecules */, ECMC_SCOPE_INTRA = 2 /* same molecule */ };
typedef struct {
    int32_t potential;   /* ECMC_POT_* */
    int32_t scope;       /* ECMC_SCOPE_* */
    int32_t use_charge;  /* multiply each pair by c_i c_j */
    int32_t reserved;
    double par[4];
} ecmc_energy_term;
int ecmc_energy_workspace_bytes(const ecmc_energy_term *term, int32_t n, int64_t *bytes);
int ecmc_energy_forward(const ecmc_energy_term *term, int32_t n, const double *d_pos, const double *d_charge,
                        const int32_t *d_molecule, double system_length, double *d_energy, void *d_work,
                        void *stream);
int ecmc_energy_backward(const ecmc_energy_term *term, int32_t n, const double *d_pos, const double *d_charge,
                         const int32_t *d_molecule, double system_length, double *d_grad_pos, double *d_grad_par,
                         double *d_grad_charge, void *d_work, void *stream);

/* ---- math and random stream (device) ----------------------------------------------------- */
/* Norms of the shortest separations of all n (n - 1) / 2 pairs i < j of n positions (n x dim
 * row-major, device memory), ordered (0,1), (0,2), ..., (1,2), ...: what
 * SeparationOutputHandler.write prints for point masses
 * (input_output_handler/output_handler/separation_output_handler.py:81-95, with
 * setting/hypercubic_setting.py:120-172 and base/vectors.py:29-43). n <= 65535. */
int ecmc_separation_norms(int32_t dim, int64_t n, const double *d_positions, double system_length, double *d_out,
                          void *stream);
int ecmc_math_log(int64_t n, const double *d_x, double *d_out, void *stream);
int ecmc_math_pow(int64_t n, const double *d_x, const double *d_y, double *d_out, void *stream);
/* Largest error over every float with bit pattern in [bits_lo, bits_hi) of the device's FP32
 * log2f (fn 0: |error| / max(|log2 x|, 1)) or exp2f (fn 1: relative error) against FP64
 * log2 / exp2, as
 * the bit pattern of a non-negative double combined into *d_out by atomic max (clear it
 * first).  No reference counterpart: it measures the error the FP32 bounds of the batched
 * kernel must cover. */
int ecmc_fp32_math_error(int32_t fn, uint32_t bits_lo, uint32_t bits_hi, unsigned long long *d_out, void *stream);
/* Draw n MT19937 words after seeding like random.seed(seed), on the device (one workgroup,
 * parallel twist). */
int ecmc_mt_words(uint64_t seed, int64_t n, uint32_t *d_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
