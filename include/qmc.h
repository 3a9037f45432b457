/* qmc.h -- thin C ABI over the MI355X Metropolis sweep engine (libqmc.so).
 *
 * This is the only door between the C++ host solvers (SA / PT / PA, JSON
 * API) and HIP.  Plain pointers and sizes; integer status codes; caller-owned
 * host buffers are copied before a call returns; device buffers are owned by
 * the handles.  A handle is used from one host thread at a time.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference's cpp/ directory):
 *   qmc_problem_create   graph::normalize_edges + Graph::populate_node_edge_ids
 *                        (graph/graph.h:74-177, 517-534) feeding IsingTermCached
 *                        (model/ising.h:305-342) / PuboWithCounter (model/pubo.h:285-399)
 *   qmc_walkers_init     Walker::init -> get_random_spin_state / get_random_binary_state
 *                        (markov/walker.h:50-58, model/ising.h:23-41, model/pubo.h:16-34)
 *   qmc_walkers_sweep    Walker::make_sweep / attempt_transition / Metropolis::accept
 *                        (markov/walker.h:144-153, 209-229; markov/metropolis.h:64-104),
 *                        the OMP replica loop of SimulatedAnnealing::make_step
 *                        (solver/simulated_annealing.h:76-110) and, with a PT ladder,
 *                        ParallelTempering::make_step (solver/parallel_tempering.h:134-189)
 *   qmc_walkers_pt       ParallelTempering::make_step with ParallelTempering::accept
 *                        and the swap loop (solver/parallel_tempering.h:126-189)
 *   qmc_batch_run        the OMP loops over replicas of many independent solves
 *                        fused into one launch (no counterpart: the reference runs
 *                        one solve per process)
 *   qmc_walkers_read     Walker::cost / get_lowest_cost / get_lowest_state
 *                        (markov/walker.h:232-305)
 *   qmc_energy           AbstractIsing::calculate_cost / PuboWithCounter::calculate_cost
 *                        (model/ising.h:94-102, model/pubo.h:303-333)
 */
#ifndef QMC_H
#define QMC_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* status codes: utils::Error values (utils/error_handling.h:25-52) */
enum {
  QMC_OK = 0,
  QMC_MEMORY_LIMITED = 1,        /* MemoryLimited: admission / allocation refused */
  QMC_DUPLICATED_VARIABLE = 101, /* DuplicatedVariable */
  QMC_INVALID_ARGUMENT = 107,    /* ValueError */
  QMC_RUNTIME_ERROR = 206,       /* UnmetPrecondition (e.g. a ladder barrier timed out) */
  QMC_DEVICE_ERROR = 401         /* no counterpart: the device or the HIP runtime failed */
};

enum { QMC_ISING = 0, QMC_PUBO = 1 };
enum { QMC_COUPLING_I32 = 0, QMC_COUPLING_F32 = 1 };
/* beta source for a sweep launch */
enum { QMC_BETA_SCHEDULE = 0, QMC_BETA_PER_REPLICA = 1 };

typedef struct qmc_ctx_s* qmc_ctx;
typedef struct qmc_problem_s* qmc_problem;
typedef struct qmc_walkers_s* qmc_walkers;
typedef struct qmc_batch_s* qmc_batch;

typedef struct {
  int32_t n_vars;
  int64_t n_terms;
  int64_t n_incidences;
  int32_t n_classes;       /* colour classes of the sweep order */
  int32_t lanes;           /* positions processed side by side in one class round */
  int32_t coupling;        /* QMC_COUPLING_* chosen from the coefficients */
  int32_t pairwise;        /* every term has at most two variables */
  int32_t max_degree;      /* max incident terms per variable */
  double mean_degree;
  int32_t max_class;       /* size of the largest colour class */
  int32_t entry_bytes;     /* bytes per incidence of the sliced rows: 4 (pairwise, integer
                              coefficients in [-128, 127], 24-bit position + int8 coefficient),
                              8 (pairwise, position + int32 or fp32 coefficient), 16 (terms of up
                              to 4 variables), 0 (incidence lists) */
} qmc_problem_info;

/* thread-local message of the last failing call */
const char* qmc_last_error(void);

int qmc_ctx_create(int device, qmc_ctx* out);
void qmc_ctx_destroy(qmc_ctx ctx);
int qmc_sync(qmc_ctx ctx);
/* Compute units of the context's device (no counterpart in the reference,
 * which sizes its work by OpenMP threads): the host solvers size default
 * launch shapes by it. */
int qmc_ctx_compute_units(qmc_ctx ctx, int32_t* compute_units);

/* Problem in internal ids 0..n_vars-1 (first-appearance order) as a term
 * CSR; ids sorted inside each term, no constant terms.  lanes = default
 * class-round width of walkers on it (0: each walker set picks its own from
 * the largest colour class and its work-group count: from 64 lanes, doubled
 * while its work-groups hold fewer than 16 waves per compute unit and one
 * round does not cover the largest class, or 8 lanes per position of it for
 * integer problems in sliced entries, whose small classes share positions
 * among lanes; at most 512 lanes x ceil(replicas_per_group / 8)). */
int qmc_problem_create(qmc_ctx ctx, int kind, int32_t n_vars, int64_t n_terms,
                       const int64_t* term_offsets, const int32_t* term_vars,
                       const double* coeffs, int32_t lanes, qmc_problem* out);
/* The same with a choice of visit order: QMC_VISIT_COLOR_CLASSES (what
 * qmc_problem_create does: greedy colouring, fewest classes) or
 * QMC_VISIT_INDEX_LEVELS (class of variable v = 1 + the largest class of a
 * term-sharing neighbour u < v: sweeps make the decisions of an index-order
 * sweep, markov/walker.h:144-153, and end in its states; the lowest state is
 * tracked along the class order). */
enum { QMC_VISIT_COLOR_CLASSES = 0, QMC_VISIT_INDEX_LEVELS = 1 };
int qmc_problem_create_visit(qmc_ctx ctx, int kind, int32_t n_vars, int64_t n_terms,
                             const int64_t* term_offsets, const int32_t* term_vars, const double* coeffs,
                             int32_t lanes, int32_t visit, qmc_problem* out);
void qmc_problem_destroy(qmc_problem p);
int qmc_problem_get_info(qmc_problem p, qmc_problem_info* info);
/* Sweep visit order (variable ids) for `lanes` positions per class round
 * (0: the problem's default) and class offsets (n_classes + 1). */
int qmc_problem_order(qmc_problem p, int32_t lanes, int32_t* order, int32_t* class_offsets);

/* R walkers on one problem; replicas_per_group (1..32) replicas share one
 * work-group.  Walkers form independent runs of replicas_per_seed
 * consecutive replicas: global replica g = first_replica + r belongs to run
 * k = g / replicas_per_seed, whose RNG key is (seed + k, stream) and whose
 * replica ids are g % replicas_per_seed -- so run k reproduces a single run
 * started with seed + k. */
int qmc_walkers_create(qmc_problem p, int32_t n_replicas, int32_t replicas_per_group,
                       uint32_t seed, uint32_t stream, int32_t first_replica,
                       int32_t replicas_per_seed, qmc_walkers* out);
void qmc_walkers_destroy(qmc_walkers w);
/* random initial states (counter contract, INIT domain), cost, lowest = cost.
 * Recorded here and performed by the next call that needs the state (for a
 * batch: by qmc_batch_run, one launch for every pending member). */
int qmc_walkers_init(qmc_walkers w, uint32_t epoch);
/* set states (n_replicas * n_vars bytes, 1 = the reference's `true` bit) */
int qmc_walkers_set_states(qmc_walkers w, const uint8_t* states);
int qmc_walkers_set_betas(qmc_walkers w, const double* betas /* n_replicas */);
/* Run n_sweeps sweeps starting at global sweep index sweep0.
 *   beta_mode QMC_BETA_SCHEDULE   : beta of sweep s = schedule[s - sweep0] (same for all replicas)
 *   beta_mode QMC_BETA_PER_REPLICA: betas set by qmc_walkers_set_betas (or a PT ladder)
 * ordered = 1 visits variables in index order (one lane per replica); 0 uses
 * the coloured order of qmc_problem_order. */
int qmc_walkers_sweep(qmc_walkers w, int64_t sweep0, int32_t n_sweeps, int beta_mode,
                      const double* schedule, int ordered);
/* Every later sweep of this walker set -- qmc_walkers_sweep and batch
 * launches -- visits variables in index order (ordered != 0; at most 64
 * replicas per group, no replica-exchange ladder) or in colour-class order
 * (0, the default).  The index order is the reference's sweep
 * (markov/walker.h:144-153) step for step; the colour classes run their
 * variables in parallel. */
int qmc_walkers_set_sweep_order(qmc_walkers w, int32_t ordered);
/* Observe the parallel tempering steps that follow (steps 0..max_steps-1
 * of their global numbering; they then run as graph-replayed sweep and
 * exchange launches): per ladder, step and ladder position k, 4 doubles --
 * the cost after the step's sweeps, the transitions accepted during them,
 * the direction label before the swaps (0 up, 1 down: it travels with the
 * states and is reset at both ends) and the swap of pair (k, k+1) (-1: not
 * attempted), the observations of parallel_tempering.h:141-188.
 * qmc_walkers_pt_log waits for the stream and copies the first n_values
 * values (layout [ladder][step][position][4]). */
int qmc_walkers_pt_observe(qmc_walkers w, int32_t ladder, int64_t max_steps);
int qmc_walkers_pt_log(qmc_walkers w, int64_t n_values, double* out);
/* Parallel tempering over ladders of `ladder` consecutive replicas with
 * betas ladder_betas[0..ladder) (descending).  Runs n_steps steps of
 * sweeps_per_replica sweeps, exchanging pairs (i,i+1), i = step mod 2, after
 * each step; steps are numbered from step0. */
int qmc_walkers_pt(qmc_walkers w, int32_t ladder, const double* ladder_betas, int64_t step0,
                   int32_t n_steps, int32_t sweeps_per_replica);
/* Plan the next sweeps of every work-group without launching (for batches):
 * next global sweep index sweep0, n_sweeps left, beta per sweep from
 * `schedule` (n_sweeps values, shared by all replicas) or, when NULL, the
 * per-replica betas. */
int qmc_walkers_plan(qmc_walkers w, int64_t sweep0, int32_t n_sweeps, const double* schedule);
/* Plan parallel tempering inside each work-group (replicas_per_group ==
 * ladder): n_steps steps of sweeps_per_replica sweeps, exchanges after each. */
int qmc_walkers_plan_pt(qmc_walkers w, int32_t ladder, const double* ladder_betas, int64_t step0,
                        int32_t n_steps, int32_t sweeps_per_replica);
/* Population resampling (solver/population.h:171-218, Population::resample):
 * replica r of the new population, r < new_R, becomes a copy of replica
 * src[r] of the current one -- state, lowest state, costs, beta and counters.
 * new_R may not exceed the replica count the walkers were created with. */
int qmc_walkers_resample(qmc_walkers w, const int32_t* src, int32_t new_R);
/* Current number of replicas. */
int qmc_walkers_count(qmc_walkers w, int32_t* n_replicas);

/* Population annealing with the population on the device
 * (solver/population_annealing.h:262-616; replaces the host loop of
 * make_step + resample, Poisson branch).  The walkers, created with the
 * capacity as their replica count, hold a population of `population`
 * citizens (replicas 0..R-1, R kept in device memory; the rest idle).
 * beta_table[k] = the schedule's beta k steps after the last restart,
 * k = 0..n_table-1 (n_table > the steps to run).  Every citizen starts at
 * beta_table[0]; initialise the states with qmc_walkers_init afterwards. */
int qmc_walkers_pa_begin(qmc_walkers w, int32_t population, double alpha, double scale,
                         const double* beta_table, int32_t n_table, uint32_t seed,
                         int32_t sweeps_per_step);
/* Enqueue one step after the step's sweeps, without waiting: global lowest
 * update, cost statistics, weights exp(-dbeta (E - E_min) scale), Poisson
 * offspring (counter RNG (citizen, step, k, 4), key (seed, 0)) with mean
 * population * w / sum(w), the copies, the family statistic and the restart
 * rule (R / rho < alpha or zero variance: fresh citizens, population topped
 * up to `population`). */
int qmc_walkers_pa_step(qmc_walkers w);
/* Stepping size and resampling method, set after qmc_walkers_pa_begin and
 * before the first step (population_annealing.h:420-457, 540-578,
 * 649-751).  strategy: QMC_PA_SCHEDULE (default: beta_table differences),
 * QMC_PA_FRICTION_TENSOR (dbeta = constant / sqrt(zeta), zeta = sigma *
 * sweeps_per_step / (ln sigma - ln |gamma|) from the costs before and after
 * the step's sweeps, sigma = sqrt(var_before var_after), gamma their
 * covariance), QMC_PA_ENERGY_VARIANCE (dbeta = constant / sqrt(sigma)) or
 * QMC_PA_CONSTANT_CULLING (dbeta = find_delta_beta of the culling estimate
 * with culling fraction `constant`); the adaptive ones need a beta_table of
 * {0, 0}.  constant_population != 0: offspring floor(population * w / sum w)
 * plus a fill to the live size drawn without replacement in proportion to
 * the residuals (utils::RandomSelector; pick j's uniform from the counter
 * (j, step, 1, 5)) instead of Poisson offspring.  exp_repopulation != 0: a
 * restart doubles the target population first (the reference's feature
 * switch 2, FEATURE_PA_EXP_REPOPULATION, population_annealing.h:463-470);
 * past the capacity the target is clamped and `overflow` is set. */
enum {
  QMC_PA_SCHEDULE = 0,
  QMC_PA_FRICTION_TENSOR = 1,
  QMC_PA_ENERGY_VARIANCE = 2,
  QMC_PA_CONSTANT_CULLING = 3
};
int qmc_walkers_pa_strategy(qmc_walkers w, int32_t strategy, double constant, int32_t constant_population,
                            int32_t exp_repopulation);
/* Per-step records of the next steps (up to max_steps; before the first
 * step): 13 doubles per step -- epoch before the step, mean cost after the
 * sweeps, population, sigma, tau, zeta (NaN where undefined), stepping size,
 * family count, sum of family sizes, size of the family with the largest
 * id, rho, restart flag, resampled flag -- the observations of
 * population_annealing.h:262-470.  qmc_walkers_pa_log waits for the stream
 * and copies the first n_steps records. */
#define QMC_PA_LOG_FIELDS 13
int qmc_walkers_pa_observe(qmc_walkers w, int32_t max_steps);
int qmc_walkers_pa_log(qmc_walkers w, int32_t n_steps, double* records);
typedef struct {
  int32_t population;  /* live citizens */
  int32_t restarts;
  int32_t overflow;    /* offspring exceeded the capacity (population clamped) */
  int32_t have_best;
  int64_t steps;
  int64_t attempts;    /* spin-flip attempts of the steps so far */
  double best;         /* lowest cost any citizen reached */
  double beta;         /* current beta */
  double rho;          /* family statistic of the last resampling */
  double delta_beta;   /* stepping size of the last step */
  int32_t target;      /* target population (doubled by restarts under exp_repopulation) */
  int32_t pad_;
  int64_t evaluations; /* cost + difference evaluations as the reference counts them */
} qmc_pa_info;
/* Waits for the stream; best_state (n_vars bytes, may be NULL) receives the
 * state of `best`. */
int qmc_walkers_pa_info(qmc_walkers w, qmc_pa_info* info, uint8_t* best_state);
/* Ends device population mode: the walkers hold the live citizens as their
 * replicas (qmc_walkers_count, _read, _energy see them). */
int qmc_walkers_pa_end(qmc_walkers w);
/* Largest number of planned sweeps not yet run by any work-group. */
int qmc_walkers_remaining(qmc_walkers w, int64_t* sweeps_left);

/* A batch launches the work-groups of many walker sets -- different
 * problems, sizes and progress -- together: one launch per kernel variant
 * (model kind x pairwise x coupling type), largest problems first. */
int qmc_batch_create(qmc_ctx ctx, qmc_batch* out);
void qmc_batch_destroy(qmc_batch b);
int qmc_batch_add(qmc_batch b, qmc_walkers w);
int qmc_batch_remove(qmc_batch b, qmc_walkers w);
int qmc_batch_clear(qmc_batch b);
/* Asynchronous: every member work-group runs up to max_sweeps planned sweeps. */
int qmc_batch_run(qmc_batch b, int32_t max_sweeps);
/* Asynchronous: every planned sweep as launches of sweeps_per_launch sweeps,
 * recorded once into a HIP graph (again only when the batch changes) and
 * replayed; the units' progress lives on the device, so replays take no
 * argument updates.  Not for split PT ladders. */
int qmc_batch_run_chunked(qmc_batch b, int32_t sweeps_per_launch);
/* Wait for the runs since the previous wait; ms = their device time span
 * (HIP events from the first run's start to the last run's end, including
 * any work enqueued between the runs). */
int qmc_batch_wait(qmc_batch b, double* ms);
/* Read back every member in the order added, in one launch and one copy per
 * array: member k owns the next R_k entries of lowest_energy (FP64 energy of
 * each replica's lowest state, recomputed from the term list as
 * qmc_walkers_energy does) and accepted, and the next R_k * n_vars_k bytes
 * of lowest_states.  Replaces the per-solver Walker::get_lowest_cost /
 * get_lowest_state reads (markov/walker.h:232-305) of many solves.  Any
 * pointer may be NULL. */
int qmc_batch_read(qmc_batch b, double* lowest_energy, uint64_t* accepted, uint8_t* lowest_states);

/* Positions processed side by side in a class round (64, 128, 256 or 512,
 * lanes * ceil(replicas_per_group / 8) <= 512; 0 = the problem default).
 * Changes the coloured visit order from the next sweep on. */
int qmc_walkers_set_lanes(qmc_walkers w, int32_t lanes);
int qmc_walkers_get_lanes(qmc_walkers w, int32_t* lanes);
/* How ladders split over several work-groups exchange: 0 (default) inside
 * one cooperative launch with a per-ladder barrier after every step; 1
 * through graph-replayed (sweep, exchange) launches. */
int qmc_walkers_set_pt_mode(qmc_walkers w, int mode);

/* Read back per replica: current cost, lowest cost, current beta, accepted
 * counts, current and lowest states (n_replicas*n_vars bytes).  Any pointer
 * may be NULL. */
int qmc_walkers_read(qmc_walkers w, double* cost, double* lowest, double* beta,
                     uint64_t* accepted, uint8_t* states, uint8_t* lowest_states);
/* One replica's current and/or lowest state (n_vars bytes each; NULL skips). */
int qmc_walkers_read_replica(qmc_walkers w, int32_t replica, uint8_t* state, uint8_t* lowest_state);
/* Recompute energies (FP64, on device) of the current or lowest states. */
int qmc_walkers_energy(qmc_walkers w, int lowest, double* energies);
/* PT ladder labels: slot holding beta index k for every ladder (n_replicas) */
int qmc_walkers_labels(qmc_walkers w, int32_t* labels);

/* Energies of arbitrary states (FP64 on device). */
int qmc_energy(qmc_problem p, int32_t n_states, const uint8_t* states, double* energies);

/* qmc_energy without allocating: no hipMalloc, hipFree or allocating library
 * call.  No counterpart in the reference; for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid it).  The states
 * and energies are staged in a fixed device array of
 * QMC_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load until the
 * process ends: n_states * n_vars (rounded up to 256) + 8 * n_states bytes
 * must fit, else QMC_INVALID_ARGUMENT.  Every call on one GPU shares that
 * array; calls hold a per-device lock until their results are copied back,
 * so concurrent callers take turns.  Synchronous, on the problem's context
 * stream; results equal qmc_energy's. */
#define QMC_NOALLOC_SCRATCH_BYTES (16u << 20)
int qmc_energy_noalloc(qmc_problem p, int32_t n_states, const uint8_t* states, double* energies);

/* Device-side delta-energy of every variable for given states (the GPU's own
 * dE arithmetic: int32 or fp32 sums in CSR order), out[n_states*n_vars]. */
int qmc_delta_all(qmc_problem p, int32_t n_states, const uint8_t* states, double* out);

/* Philox4x32-10 words computed on the device: out[4*n] for counters ctr[4*n]. */
int qmc_philox(qmc_ctx ctx, int32_t n, const uint32_t* ctr, const uint32_t key[2], uint32_t* out);
/* Aggregate LDS read bandwidth of the device in bytes/s, measured by a
 * kernel that streams 16-byte LDS reads on every CU (no counterpart in the
 * reference: it calibrates the LDS peak the roofline figures are quoted
 * against). */
int qmc_lds_bandwidth(qmc_ctx ctx, double* bytes_per_s);

/* Timing of the last sweep launch on this walker set (HIP events on the
 * stream it ran on), milliseconds. */
int qmc_walkers_last_ms(qmc_walkers w, double* ms);

#ifdef __cplusplus
}
#endif
#endif
