// Host solvers over the device sweep engine, keeping the reference's
// lifecycle configure(json) -> init() -> run() -> finalize() -> get_result()
// (solver/solver.h:37-50,147-156,224) and parameter names (Appendix B of the
// reference docs: solver.h:166-217, stepping_solver.h:55-79,
// simulated_annealing.h:123-209, parallel_tempering.h:192-234,
// population_annealing.h:632-778).
#pragma once
#include <limits>
#include <memory>
#include <optional>
#include <set>
#include <string>
#include <vector>

#include "json.h"
#include "model.h"
#include "observer.h"
#include "schedule.h"

namespace qiotk {

struct Timing {
  double device_ms = 0.0;  // summed HIP-event time of sweep launches
  int64_t launches = 0;
};

class Solver {
 public:
  virtual ~Solver() = default;
  void set_model(const Model* m, DeviceModel* dm, qmc_ctx ctx) {
    model_ = m;
    dmodel_ = dm;
    ctx_ = ctx;
  }
  virtual std::string target() const = 0;
  // {"params": {...}} as the reference's Solver::configure takes it
  virtual void configure(const Json& doc);
  virtual void init() = 0;
  virtual void run() = 0;
  virtual void finalize() = 0;
  // Runner::get_run_output (app/runner.cc:518-537): init(); when it fails
  // with MemoryLimited and the memory-saving form is not on yet, switch it on
  // (FEATURE_USE_MEMORY_SAVING, utils/config.h:58) and init again.
  void init_with_retry();
  // The memory-saving form: at most 8 replicas per work-group (one byte
  // lane, the smallest LDS footprint per position) and, for population
  // annealing, room for the population to grow by half instead of 4x.
  // Switched on by "enabled_features": [3] or by the retry above.
  bool memory_saving() const { return memory_saving_; }
  void set_memory_saving(bool on);
  // Fused launches with other solves: plan the whole run on the device and
  // add the walkers to `b`; false when this solve needs host-side steps.
  virtual bool batch_plan(qmc_batch) { return false; }
  // Replicas this solve added to a batch; their results come back from
  // qmc_batch_read in the order the solves were added (0: none added).
  int64_t batch_replicas() const { return batched_replicas_; }
  // Replicas of one independent run as configured (SA restarts, PT ladder
  // length, PA population); known after configure().
  virtual int64_t run_replicas() const { return 1; }
  // replicas the solve allocates on the device (population annealing
  // reserves room for its population to fluctuate or grow)
  virtual int64_t device_replicas() const { return run_replicas(); }
  // Serving form: every round the server calls serve_plan() to add this
  // solve's next device work to the shared batch (false: nothing to add),
  // runs the batch, then serve_advance(), which enqueues what follows the
  // round on the device without waiting.  Several rounds may be enqueued
  // before the server waits for the device and asks serve_done() whether
  // the solve is complete (finalize() next).
  virtual bool serve_plan(qmc_batch b);
  virtual void serve_advance() {}
  virtual bool serve_done();
  // the solve's walker set on the device (null before init or when empty)
  virtual qmc_walkers walkers() const { return nullptr; }
  // Sweeps per device launch when fused into a batch (0: one persistent launch).
  int64_t sweeps_per_launch() const { return sweeps_per_launch_; }
  // finalize() from this solve's slices of a batched read-back: FP64 energy
  // of each replica's lowest state, accepted counts, lowest states.
  void finalize_from(const double* lowest_energy, const uint64_t* accepted, const uint8_t* states);
  // Best cost (constant terms included) and its configuration in the
  // reference's rendering (Ising +1/-1, PUBO 1/0) by internal variable id;
  // false for an empty model.
  bool best(double* cost, int8_t* config) const;
  // Resolved run plan (steps, betas), computable without a device.
  virtual Json describe() const;
  virtual Json get_result() const;
  Json get_solutions() const;  // ModelSolver::get_solutions (model_solver.h:111-150)
  // the serving loop can advance this solve round by round
  virtual bool servable() const { return true; }

  int64_t attempts() const { return attempts_; }
  int64_t accepted() const { return accepted_; }
  const Timing& timing() const { return timing_; }
  double lowest_cost() const { return lowest_cost_; }
  uint32_t seed() const { return seed_; }
  // the `threads` parameter (0: not given)
  int thread_count() const { return threads_; }

 protected:
  // SteppingSolver (stepping_solver.h:112-268): fewer than steps_per_tick steps
  // are never run, because the step limit is checked after each tick.
  int64_t steps_to_run() const;
  // Steps a run with an evaluation limit takes (SteppingSolver::run,
  // stepping_solver.h:119-245): ticks of steps_per_tick steps, growing to
  // max(n + 1, 1.2 n) after every tick that takes under 0.95 tick_every (all
  // of them here), the last clamped to the step limit; after a tick the step
  // limit is checked, then evaluations(steps so far) against eval_limit,
  // which stops the run with the reference's message.  Without a limit:
  // steps_to_run().
  template <class F>
  int64_t eval_limited_steps(F evaluations) {
    const int64_t total = steps_to_run();
    if (!eval_limit_) return total;
    int64_t s = 0, tick = steps_per_tick_, next = tick;
    while (true) {
      s += next;
      tick = (int64_t)std::max((double)tick + 1.0, (double)tick * 1.2);
      if (s >= total) return total;
      next = std::min(total - s, tick);
      const int64_t e = evaluations(s);
      if (e >= *eval_limit_) {
        exit_reason_ = "Stop due to Evaluation limit: " + std::to_string(*eval_limit_) +
                       ", evaluations:" + std::to_string(e);
        return s;
      }
    }
  }
  bool has_time_limit() const { return (bool)time_limit_; }
  // time_limit and cost_limit after `steps` steps of a run started at
  // wall-clock t0 (seconds), the lowest cost so far given when has_low
  // (stepping_solver.h:151-152, 195-258); sets exit_reason_ in the
  // reference's words when one is reached
  bool limits_reached(int64_t steps, double t0, bool has_low, double low);
  static double now_s();
  void record(double cost, const std::vector<uint8_t>& state);
  // ModelSolver::populate_solutions: best `number_of_solutions` walkers
  void populate(const std::vector<double>& lowest, const std::vector<uint8_t>& states, int64_t R);
  void populate(const double* lowest, const uint8_t* states, int64_t R);
  void set_output(const std::string& k, Json v) { out_params_[k] = std::move(v); }

  const Model* model_ = nullptr;
  DeviceModel* dmodel_ = nullptr;
  qmc_ctx ctx_ = nullptr;
  uint32_t seed_ = 0;
  int64_t step_limit_ = 10000;
  int64_t steps_per_tick_ = 10;
  std::optional<double> time_limit_;
  std::optional<double> cost_limit_;
  std::optional<int64_t> eval_limit_;
  int number_of_solutions_ = 1;
  int threads_ = 0;
  int lanes_ = 0;  // positions per class round on the device (0: automatic)
  int positions_per_lane_ = 0;  // else lanes from the largest class (0: library default)
  int64_t sweeps_per_launch_ = 0;
  // "sweep_order": "colored" 0 (greedy colour classes), "index" 1 (index
  // levels: the decisions and end-of-sweep states of the reference's
  // index-order sweep, run class by class), "sequential" 2 (one lane per
  // replica walking the index order), absent -1 (the solver's default)
  int sweep_order_ = -1;
  // feature switches of utils/config.h:50-62 named by "enabled_features" /
  // "disabled_features" (solver.h:205-217); only 2, FEATURE_PA_EXP_REPOPULATION,
  // changes what this engine does
  std::set<int> enabled_features_;
  // the reference's observer (observe/observer.h) and the "step" observation
  // SteppingSolver::run makes before every step (stepping_solver.h:127-133);
  // per_step(s) adds the solver's own observations of step s
  Observer observer_;
  // replays the run's observations into observer_ (end of finalize)
  virtual void observe_run() { observe_steps(steps_done_, [](int64_t) {}); }
  template <class F>
  void observe_steps(int64_t steps, F per_step) {
    if (observer_.empty()) return;
    for (int64_t s = 0; s < steps; ++s) {
      observer_.set_label("step", s);
      observer_.observe("step", (double)s);
      per_step(s);
      observer_.clear_label("step");
    }
  }
  bool feature_enabled(int f) const { return enabled_features_.count(f) > 0; }
  bool memory_saving_ = false;
  // drops what init() created on the device (before a retry)
  virtual void release_device() {}
  // replicas per work-group in the memory-saving form: at most 8, and a
  // divisor of `divide` when it is positive
  int saving_group(int B, int divide = 0) const;
  virtual bool index_by_default() const { return false; }
  int visit() const {
    return sweep_order_ == 1 || (sweep_order_ < 0 && index_by_default()) ? QMC_VISIT_INDEX_LEVELS
                                                                         : QMC_VISIT_COLOR_CLASSES;
  }
  bool sequential() const { return sweep_order_ == 2; }
  int lanes_for(int replicas_per_group) const;
  int compute_units() const;  // of the context's device (default launch shapes)
  void apply_lanes(qmc_walkers w, int replicas_per_group) const;
  int64_t steps_done_ = 0;
  int64_t attempts_ = 0, accepted_ = 0;
  Timing timing_;
  bool have_lowest_ = false;
  double lowest_cost_ = 0.0;
  std::vector<uint8_t> lowest_state_;
  std::vector<double> lowest_costs_;
  std::vector<std::vector<uint8_t>> lowest_states_;
  Json out_params_ = Json::object();
  std::string exit_reason_;
  int64_t batched_replicas_ = 0;
  bool served_ = false;
  // served round by round (serve_plan): a whole-run plan with time or cost
  // limits checked by serve_done() between waits for the device
  bool serving_ = false;
  double serve_t0_ = 0.0;
  int64_t sweeps_per_step_ = 1;
  friend class ParameterFree;  // reads its workers' results
};

// simulated_annealing.h
class SimulatedAnnealing : public Solver {
 public:
  std::string target() const override { return "simulatedannealing.qiotoolkit"; }
  void configure(const Json& doc) override;
  void init() override;
  void run() override;
  void finalize() override;
  bool batch_plan(qmc_batch b) override;
  Json describe() const override;
  int restarts() const { return restarts_; }
  int64_t run_replicas() const override { return restarts_; }
  qmc_walkers walkers() const override { return w_; }

  int64_t planned_steps();

 protected:
  void release_device() override;

 private:
  bool use_inverse_temperature_ = false;
  int restarts_ = 64;
  int group_ = 0;
  Schedule schedule_;
  qmc_walkers w_ = nullptr;
  // one beta per step, shared by the solves of one schedule (read-only)
  std::shared_ptr<const std::vector<double>> betas_;
  std::vector<double> schedule_betas() const;
  std::vector<double> lowest_;
  std::vector<uint8_t> lowest_states_flat_;

 public:
  ~SimulatedAnnealing() override;
};

// parallel_tempering.h
class ParallelTempering : public Solver {
 public:
  std::string target() const override { return "paralleltempering.qiotoolkit"; }
  void configure(const Json& doc) override;
  void init() override;
  void run() override;
  void finalize() override;
  bool batch_plan(qmc_batch b) override;
  Json describe() const override;
  ~ParallelTempering() override;
  std::vector<double> ladder() const;
  int64_t run_replicas() const override { return (int64_t)ladder().size(); }
  qmc_walkers walkers() const override { return w_; }

 protected:
  void observe_run() override;
  void release_device() override;

  int64_t planned_steps();

 private:
  bool observe_dynamics_ = false;  // the observer watches acc_rate, avg_cost, avg_dir or swap_rate
  bool use_inverse_temperatures_ = true;
  Schedule temperatures_;
  int sweeps_per_replica_ = 1;
  int group_ = 0;
  std::vector<double> betas_;  // descending
  qmc_walkers w_ = nullptr;
  std::vector<double> lowest_;
  std::vector<uint8_t> lowest_states_flat_;
};

// population_annealing.h: device sweeps and device resampling (Poisson),
// linear or geometric beta schedule.
class PopulationAnnealing : public Solver {
 public:
  std::string target() const override { return "populationannealing.cpu"; }
  void configure(const Json& doc) override;
  void init() override;
  void run() override;
  void finalize() override;
  bool serve_plan(qmc_batch b) override;
  void serve_advance() override;
  bool serve_done() override;
  qmc_walkers walkers() const override;
  int64_t run_replicas() const override { return target_population_; }
  int64_t device_replicas() const override;
  ~PopulationAnnealing() override;

 protected:
  void observe_run() override;
  // under colour classes the best energies of population annealing are
  // distributed differently from the reference's (a 12x12 lattice at
  // population 64 finds lower energies with fewer restarts)
  bool index_by_default() const override { return true; }
  void release_device() override { impl_.reset(); }

 private:
  void finish_step();
  int64_t step_ = 0, sweep_ = 0;
  bool step_planned_ = false;  // serving: this round's sweeps are planned, its step not enqueued
  bool limit_hit_ = false;     // serving: a time or cost limit ended the run
  bool in_batch_ = false;
  double alpha_ = 2.0;
  int64_t target_population_ = 64;
  bool constant_population_ = false;
  std::string strategy_ = "linear_schedule";
  int strategy_id_ = QMC_PA_SCHEDULE;
  double strategy_const_ = 0.0;  // friction constant or culling fraction
  Schedule beta_;
  double beta_start_ = 0.0, beta_stop_ = 5.0;
  int sweeps_per_replica_ = 1;
  double cur_beta_ = 0.0;
  int64_t restarts_ = 0;
  int group_ = 0;
  std::shared_ptr<void> impl_;
};

// Parameter-free simulated annealing or parallel tempering
// (solver/parameter_free_solver.h:11-257 with the linear search of
// parameter_free_linear_solver.h:8-30 and strategy/linear_search.cc:13-122;
// adapters sa_parameter_free.h:18-175, pt_parameter_free.h:18-190): repeated
// worker runs whose sweeps walk a set of sample points that stretches (best
// at the last point) or shrinks (best at the first) by 1.5 after every pass,
// until 8 runs in a row bring no lower cost or the time limit (5 s unless
// given) is spent.  Every pass over the sample points is planned before its
// first run, so its runs are fused into one device batch; their results are
// then taken in order with the reference's bookkeeping, and the runs after a
// stop inside a pass are discarded.
class ParameterFree : public Solver {
 public:
  explicit ParameterFree(bool tempering) : tempering_(tempering) {}
  std::string target() const override;
  void configure(const Json& doc) override;
  void init() override {}
  void run() override;
  void finalize() override;
  Json describe() const override;
  Json get_result() const override;
  bool servable() const override { return false; }
  int64_t run_replicas() const override { return min_num_; }

 private:
  std::unique_ptr<Solver> worker(const std::vector<double>& params, uint32_t seed) const;
  Json worker_params(const std::vector<double>& params, uint32_t seed) const;
  Json output_for(const std::vector<double>& params) const;
  double execution_cost(const std::vector<double>& params) const;
  std::vector<double> betas() const;  // PTParameterFree::determine_all_betas_

  bool tempering_;
  Json base_ = Json::object();  // the request's params, forwarded to the workers where they apply
  double max_prob_ = 0.5, min_prob_ = 0.0001, adjustment_ = 0.001;
  double beta_start_ = 0.01, beta_stop_ = 10.0;
  double max_diff_ = 0.0, min_diff_ = 0.0;
  std::vector<int> points_{10, 20, 30, 50};
  int64_t min_num_ = 64;  // min_num_restarts (SA) or min_num_betas (PT)
  int nonimprovement_limit_ = 8;
  double change_factor_ = 1.5;
  std::vector<double> best_params_;
  std::unique_ptr<Solver> last_;  // the last worker taken, source of the extra solutions
  Json training_ = Json::object();
  int64_t runs_ = 0, passes_ = 0, discarded_ = 0;
};

// Replicas per work-group when the params do not say: small groups while
// that still leaves at least 128 work-groups, so a launch fills the device.
int default_group(int64_t replicas, int compute_units);

std::unique_ptr<Solver> create_solver(const std::string& target);

}  // namespace qiotk
