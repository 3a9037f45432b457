/* CPU oracle for the QIOToolkit annealing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C++, the
 * reference's Metropolis sweep and the solvers that drive it, so that the
 * HIP product path can be checked against it.  Only tests/, the smoke()
 * entry point and bench.py's CPU-baseline leg load it; the product library
 * (qiotoolkit_amd) never links or calls it.
 *
 * Two RNG modes:
 *   ORC_RNG_REFERENCE : the reference's own streams -- pcg32 (PCG-XSH-RR
 *                       64/32) per replica, forked from a master PCG(seed)
 *                       exactly as utils::PCG::fork does
 *                       (cpp/utils/random_generator.cc:37-58).
 *   ORC_RNG_COUNTER   : the counter-based contract the GPU implements
 *                       (Philox4x32-10 keyed by seed, counter by
 *                       (variable, sweep, replica, domain)).  Same sweep,
 *                       accept, apply and best-tracking code; only the source
 *                       of the 32-bit words differs.
 */
#ifndef QIOTK_ORACLE_H
#define QIOTK_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_RNG_REFERENCE = 0, ORC_RNG_COUNTER = 1 };
enum { ORC_ARITH_F64 = 0, ORC_ARITH_F32 = 1 };
enum { ORC_ISING = 0, ORC_PUBO = 1 };

typedef struct {
  int rng_mode;          /* ORC_RNG_* */
  int arith;             /* ORC_ARITH_F64: double dE as the reference; F32: float sum in CSR order */
  int n_order;           /* 0: ordered sweep i = 0..N-1; else N = length of `order` */
  const int32_t* order;  /* visit order for one sweep (a permutation of 0..N-1) */
  uint32_t seed;         /* solver seed (reference `seed` param) */
  uint32_t stream;       /* Philox key word 1 (counter mode only) */
  int threads;           /* OpenMP threads for replica loops (0 = runtime default) */
  int64_t steps_per_tick;/* reference `steps_per_tick` (default 10) */
} orc_opts;

/* ---- RNG (cpp/utils/random_generator.cc) ---- */
void orc_pcg_uint32(int has_seed, uint32_t seed, int n, uint32_t* out);
void orc_pcg_uniform(int has_seed, uint32_t seed, int n, double* out);
/* master PCG(seed) forks n_forks children; then n uniforms from the master and
 * from each fork: out[(k)*n + t], k = 0 master, 1..n_forks forks. */
void orc_pcg_fork_uniform(uint32_t seed, int n_forks, int n, double* out);
void orc_pcg_poisson(uint32_t seed, double mean, int n, uint32_t* out);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- Metropolis accept with the ln lookup table (markov/metropolis.h:64-104) ---- */
void orc_uniform_ln(double* out /* 4*256 */);
int orc_accept(double cost_diff, double beta, const uint32_t* draws /* up to 4 */, int* n_used);

/* ---- schedules (cpp/schedule) ---- */
/* type 0 linear, 1 geometric, 2 constant; start/stop/count < 0 means unset */
double orc_schedule_value(int type, double initial, double final_, double start, double stop,
                          double x);
int orc_schedule_discretized(int type, double initial, double final_, int count, double* out);

/* ---- problems (graph/graph.h normalize_edges; model/ising.h, model/pubo.h) ---- */
void* orc_problem_new(int kind, int64_t n_terms, const double* c, const int64_t* offsets,
                      const int32_t* ids, char* err, int errlen);
void orc_problem_free(void* p);
void orc_problem_dims(void* p, int64_t* n_vars, int64_t* n_terms, int64_t* n_incidences,
                      double* const_cost, int* max_locality);
void orc_problem_names(void* p, int32_t* names);
void orc_problem_edges(void* p, int64_t* offsets, int32_t* ids, double* c);
/* spins: one byte per internal variable, 1 = the reference's `true` bit
 * (Ising s = -1, PUBO x = 0). */
double orc_cost(void* p, const uint8_t* spins);
double orc_delta(void* p, const uint8_t* spins, int64_t i, int arith);
/* 1 if no two variables of one colour class share a term */
int orc_check_classes(void* p, const int32_t* order, int n_classes, const int32_t* class_offsets);
/* initial random state of replica r (reference: get_random_spin_state from a
 * forked PCG; counter mode: Philox INIT domain). */
void orc_initial_state(void* p, const orc_opts* o, int32_t replica, uint32_t epoch, uint8_t* spins);

/* Run n_sweeps sweeps on R replicas from given states (counter mode only):
 * betas[k*R + r] for sweep sweep0+k; replica_ids[r] keys the RNG.
 * states (R*N) and costs/lowest/lowest_states are updated in place. */
int orc_run_sweeps(void* p, const orc_opts* o, int R, const int32_t* replica_ids, int64_t sweep0,
                   int n_sweeps, const double* betas, uint8_t* states, double* costs,
                   double* lowest, uint8_t* lowest_states, uint64_t* accepted);

/* ---- solvers ---- */
/* SimulatedAnnealing (solver/simulated_annealing.h).  sched_type: 0 linear
 * (beta_start/beta_stop form when use_inverse=1), 1 geometric.  Outputs per
 * restart: lowest cost/state and final cost/state. */
int orc_sa(void* p, const orc_opts* o, int sched_type, double initial, double final_,
           int use_inverse_temperature, int64_t step_limit, int restarts, double* lowest,
           uint8_t* lowest_states, double* final_costs, uint8_t* final_states, int64_t* steps_run);

/* ParallelTempering (solver/parallel_tempering.h).  values = temperatures
 * (use_inverse=0) or all_betas (use_inverse=1), unsorted as given.
 * Reference mode swaps states between beta slots; counter mode keeps states
 * in place and swaps beta labels (slot r keeps RNG key r).  Outputs per slot. */
int orc_pt(void* p, const orc_opts* o, const double* values, int n_values, int use_inverse,
           int64_t step_limit, int sweeps_per_replica, double* slot_beta, double* lowest,
           uint8_t* lowest_states, double* final_costs, uint8_t* final_states,
           double* swap_rate, int64_t* steps_run, double* acc_rate, double* avg_cost, double* avg_dir);

/* PopulationAnnealing (solver/population_annealing.h), linear/geometric
 * schedule, Poisson (variable population) or constant_population resampling.
 * Reference mode only.  Returns best cost, best state, final population size. */
void orc_poisson_counter(double lam, uint32_t seed, uint32_t step, int n, uint32_t* out);
int orc_pa_step_counter(int R, const double* cost, const double* lowest, const int32_t* family,
                        int64_t step, int64_t base_step, const double* beta_table, int n_table, int target,
                        int cap, double alpha, double scale, uint32_t seed, int have_best, double best,
                        int32_t* src, int32_t* fam_out, int32_t* iout, double* dout, const double* cost_before,
                        int strategy, double strategy_const, int constant_population, int sweeps_per_step,
                        int exp_repopulation);
int orc_toy_sa(uint32_t seed, int64_t step_limit, int restarts, double beta_start, double beta_stop,
               double* ms_steps, double* ms_values, int cap);
int orc_toy_pt(uint32_t seed, int64_t step_limit, int n, const double* temps, int sweeps_per_replica,
               double* T_out, double* acc_rate, double* avg_cost, double* avg_dir, double* swap_rate,
               double* ms_steps, double* ms_values, int cap);
int orc_selector_draw(int n, const double* weights, int k, const double* uniforms, int64_t* out);
double orc_adaptive_delta_beta(int strategy, double c, int R, const double* before, const double* after,
                               int target, double scale, int sweeps);
double orc_culling_estimate(int R, const double* shifted, double epsilon0, int target, double scale,
                            double db);
int orc_pa(void* p, const orc_opts* o, int geometric, double beta_start, double beta_stop,
           int64_t step_limit, int population, int sweeps_per_replica, double alpha,
           int constant_population, double* best_cost, uint8_t* best_state,
           int64_t* final_population, int64_t* restarts, int strategy, double strategy_const,
           int exp_repopulation);

#ifdef __cplusplus
}
#endif
#endif
