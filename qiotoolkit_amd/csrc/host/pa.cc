// PopulationAnnealing (solver/population_annealing.h:102-616) over the device
// engine: each step sweeps every citizen, then computes the stepping size
// (beta schedule, friction tensor, energy variance or constant culling), the
// Boltzmann weights, offspring counts (Poisson, or residual resampling for a
// constant population), the resampled population
// (Population::resample, solver/population.h:171-218) and restarts on the
// device, with the population size in device memory -- the host enqueues
// steps and reads the outcome once at the end.  Random numbers follow the
// counter contract: citizens' sweeps are keyed by their slot and the global
// sweep index; offspring counts by (citizen, step).
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <sstream>

#include "solver.h"

namespace qiotk {

namespace {
// the reference's matcher wording (matcher/greater-than.h:30-40): values as
// an ostream prints them
template <class T>
[[noreturn]] void not_greater(const char* name, T bound, T found) {
  std::ostringstream o;
  o << "parameter `" << name << "`: must be greater than " << bound << ", found " << found;
  throw Error(kValueError, o.str());
}

const Json& params_of(const Json& doc) {
  if (!doc.is_object() || !doc.has("params"))
    throw Error(kMissingInput, "Input field `params` is required.");
  return doc.at("params");
}
}  // namespace

struct PaState {
  qmc_walkers w = nullptr;
  int32_t cap = 0;
  ~PaState() {
    if (w) qmc_walkers_destroy(w);
  }
};

static PaState& pa_state(std::shared_ptr<void>& slot) {
  if (!slot) slot = std::make_shared<PaState>();
  return *static_cast<PaState*>(slot.get());
}

void PopulationAnnealing::configure(const Json& doc) {
  Solver::configure(doc);
  const Json& p = params_of(doc);
  if (p.has("alpha")) alpha_ = p.at("alpha").as_double();
  if (!(alpha_ > 1.0)) not_greater("alpha", 1.0, alpha_);
  target_population_ = threads_ > 0 ? threads_ : 64;
  if (p.has("population")) target_population_ = p.at("population").as_int();
  if (target_population_ <= 0) not_greater<int64_t>("population", 0, target_population_);
  if (p.has("constant_population")) constant_population_ = p.at("constant_population").as_bool();
  if (p.has("resampling_strategy")) strategy_ = p.at("resampling_strategy").as_string();
  if (p.has("sweeps_per_replica")) sweeps_per_replica_ = (int)p.at("sweeps_per_replica").as_int();
  if (sweeps_per_replica_ <= 0) not_greater<int64_t>("sweeps_per_replica", 0, sweeps_per_replica_);
  if (p.has("replicas_per_group")) group_ = (int)p.at("replicas_per_group").as_int();
  set_output("population", Json(target_population_));
  set_output("alpha", alpha_);
  set_output("constant_population", Json(constant_population_));
  set_output("sweeps_per_replica", Json((int64_t)sweeps_per_replica_));
  if (strategy_ != "linear_schedule" && strategy_ != "geometric_schedule") {
    // the adaptive stepping sizes start at beta 0 (population_annealing.h:730-774)
    auto positive = [&](const char* name, double dflt) {
      const double v = p.has(name) ? p.at(name).as_double() : dflt;
      if (!(v > 0.0)) not_greater(name, 0.0, v);
      set_output(name, v);
      return v;
    };
    if (strategy_ == "friction_tensor") {
      strategy_id_ = QMC_PA_FRICTION_TENSOR;
      strategy_const_ = positive("friction_tensor_constant", 1.0);
    } else if (strategy_ == "energy_variance") {
      strategy_id_ = QMC_PA_ENERGY_VARIANCE;
      strategy_const_ = positive("initial_culling_fraction", 0.5);
    } else if (strategy_ == "constant_culling") {
      strategy_id_ = QMC_PA_CONSTANT_CULLING;
      strategy_const_ = positive("culling_fraction", 0.1);
    } else {
      throw Error(kInitialConfig,
                  "`resampling strategy` must be one of ['linear_schedule',  'geometric_schedule', "
                  "'friction_tensor', 'energy_variance', 'constant_culling'].");
    }
    beta_start_ = 0.0;
    return;
  }
  strategy_id_ = QMC_PA_SCHEDULE;
  const bool geometric = strategy_ == "geometric_schedule";
  if (p.has("beta")) {
    beta_.configure(p.at("beta"));
  } else {
    beta_start_ = p.has("beta_start") ? p.at("beta_start").as_double() : (geometric ? 0.01 : 0.0);
    beta_stop_ = p.has("beta_stop") ? p.at("beta_stop").as_double() : 5.0;
    if (beta_stop_ < beta_start_) throw Error(kValueError, "parameter `beta_stop` must be >= `beta_start`");
    if (geometric) {
      if (!(beta_start_ > 0) || !(beta_stop_ > 0))
        throw Error(kValueError, "beta_start and beta_stop must be positive for geometric");
      beta_.make_geometric(beta_start_, beta_stop_);
    } else {
      if (beta_start_ < 0) throw Error(kValueError, "parameter `beta_start` must be >= 0");
      beta_.make_linear(beta_start_, beta_stop_);
    }
    set_output("beta_start", beta_start_);
    set_output("beta_stop", beta_stop_);
  }
  // stretch to the step limit (population_annealing.h:709-725)
  if (beta_.get_count() == 1) {
    if (step_limit_ > 1) beta_.set_stop((int)std::max<int64_t>(1, step_limit_ - 1));
    else throw Error(kMissingInput, "Either a step_limit or schedule.count must be provided.");
  }
  if (beta_.get_initial() >= beta_.get_final())
    throw Error(kValueError, "the schedule for beta must be increasing (i.e., beta.initial < beta.final).");
  beta_start_ = beta_.get_value(0);
}

PopulationAnnealing::~PopulationAnnealing() = default;

// Poisson resampling keeps the population near the target; exponential
// repopulation (feature 2) doubles the target at restarts, so it gets more
// room (past it the solve ends with MemoryLimited).  The memory-saving form
// leaves room for half the population more (Poisson fluctuations of a
// population of P are about sqrt(P)).
int64_t PopulationAnnealing::device_replicas() const {
  const int64_t P = target_population_;
  if (feature_enabled(2))
    return memory_saving_ ? std::max<int64_t>(16 * P, P + 1024) : std::max<int64_t>(64 * P, P + 4096);
  return memory_saving_ ? std::max<int64_t>(P + P / 2, P + 256) : std::max<int64_t>(4 * P, P + 256);
}

void PopulationAnnealing::init() {
  if (model_->empty()) return;
  PaState& st = pa_state(impl_);
  const bool exp_repop = feature_enabled(2);
  st.cap = (int32_t)device_replicas();
  const bool seq = sequential();
  const int B = saving_group(group_ > 0 ? group_ : seq ? 32 : default_group(target_population_, compute_units()));
  device_check(qmc_walkers_create(dmodel_->handle(visit()), st.cap, B, seed_, 0, 0, st.cap, &st.w),
               "qmc_walkers_create");
  apply_lanes(st.w, B);
  if (seq) device_check(qmc_walkers_set_sweep_order(st.w, 1), "qmc_walkers_set_sweep_order");
  // beta k steps after the last restart, k = 0..steps (get_value clamps past
  // the schedule's end); the adaptive stepping sizes start from beta 0
  std::vector<double> table(2, 0.0);
  if (strategy_id_ == QMC_PA_SCHEDULE) {
    table.resize((size_t)steps_to_run() + 2);
    for (size_t k = 0; k < table.size(); ++k) table[k] = beta_.get_value((double)k);
  }
  device_check(qmc_walkers_pa_begin(st.w, (int32_t)target_population_, alpha_, model_->scale_factor(),
                                    table.data(), (int32_t)table.size(), seed_, sweeps_per_replica_),
               "qmc_walkers_pa_begin");
  if (!observer_.empty())
    device_check(qmc_walkers_pa_observe(st.w, (int32_t)steps_to_run()), "qmc_walkers_pa_observe");
  if (strategy_id_ != QMC_PA_SCHEDULE || constant_population_ || exp_repop)
    device_check(qmc_walkers_pa_strategy(st.w, strategy_id_, strategy_const_, constant_population_ ? 1 : 0,
                                         exp_repop ? 1 : 0),
                 "qmc_walkers_pa_strategy");
  device_check(qmc_walkers_init(st.w, 0), "qmc_walkers_init");
  cur_beta_ = beta_start_;
}

void PopulationAnnealing::run() {
  if (model_->empty()) return;
  PaState& st = pa_state(impl_);
  auto run_to = [&](int64_t s) {
    while (step_ < s) {
      device_check(qmc_walkers_sweep(st.w, sweep_, sweeps_per_replica_, QMC_BETA_PER_REPLICA, nullptr, 0),
                   "qmc_walkers_sweep");
      double ms = 0;
      qmc_walkers_last_ms(st.w, &ms);
      timing_.device_ms += ms;
      timing_.launches++;
      finish_step();
    }
  };
  if (time_limit_ || cost_limit_) {
    // Ticks of steps with the limits checked in between, as SteppingSolver::run
    // does (stepping_solver.h:119-258): the step limit, then the evaluation
    // limit, the cost limit against the population's global lowest, the time
    // limit; the tick grows to max(n + 1, 1.2 n).
    const double t0 = now_s();
    const int64_t total = steps_to_run();
    int64_t tick = steps_per_tick_;
    while (step_ < total) {
      run_to(std::min(total, step_ + tick));
      tick = (int64_t)std::max((double)tick + 1.0, (double)tick * 1.2);
      if (step_ >= total) break;
      qmc_pa_info info{};
      device_check(qmc_walkers_pa_info(st.w, &info, nullptr), "qmc_walkers_pa_info");
      if (eval_limit_ && info.evaluations >= *eval_limit_) {
        exit_reason_ = "Stop due to Evaluation limit: " + std::to_string(*eval_limit_) + ", evaluations:" +
                       std::to_string(info.evaluations);
        break;
      }
      if (limits_reached(step_, t0, info.have_best != 0, info.best)) break;
    }
    return;
  }
  // with an evaluation limit the population's count is read at each tick
  const int64_t steps = eval_limited_steps([&](int64_t s) {
    run_to(s);
    qmc_pa_info info{};
    device_check(qmc_walkers_pa_info(st.w, &info, nullptr), "qmc_walkers_pa_info");
    return (int64_t)info.evaluations;
  });
  run_to(steps);
}

bool PopulationAnnealing::serve_plan(qmc_batch b) {
  if (model_->empty() || step_ >= steps_to_run() || limit_hit_) return false;
  if (!serving_) {
    serving_ = true;
    serve_t0_ = now_s();
  }
  PaState& st = pa_state(impl_);
  device_check(qmc_walkers_plan(st.w, sweep_, sweeps_per_replica_, nullptr), "qmc_walkers_plan");
  if (!in_batch_) {
    device_check(qmc_batch_add(b, st.w), "qmc_batch_add");
    in_batch_ = true;
  }
  step_planned_ = true;
  return true;
}

void PopulationAnnealing::serve_advance() {
  if (!step_planned_) return;
  step_planned_ = false;
  finish_step();
}

// Served with time or cost limits: checked after every wait for the device.
bool PopulationAnnealing::serve_done() {
  if (model_->empty() || step_ >= steps_to_run() || limit_hit_) return true;
  if (!(time_limit_ || cost_limit_)) return false;
  qmc_pa_info info{};
  device_check(qmc_walkers_pa_info(pa_state(impl_).w, &info, nullptr), "qmc_walkers_pa_info");
  limit_hit_ = limits_reached(step_, serve_t0_, info.have_best != 0, info.best);
  return limit_hit_;
}

qmc_walkers PopulationAnnealing::walkers() const {
  return impl_ ? static_cast<PaState*>(impl_.get())->w : nullptr;
}

// The rest of a step after its sweeps (population_annealing.h:262-488):
// enqueued on the device, nothing waits for it.
void PopulationAnnealing::finish_step() {
  PaState& st = pa_state(impl_);
  device_check(qmc_walkers_pa_step(st.w), "qmc_walkers_pa_step");
  sweep_ += sweeps_per_replica_;
  step_ += 1;
  steps_done_ = step_;
}

void PopulationAnnealing::finalize() {
  if (model_->empty()) return;
  PaState& st = pa_state(impl_);
  qmc_pa_info info{};
  std::vector<uint8_t> best_state((size_t)model_->n_vars);
  device_check(qmc_walkers_pa_info(st.w, &info, best_state.data()), "qmc_walkers_pa_info");
  if (info.overflow) throw Error(kMemoryLimited, "population grew beyond the reserved capacity");
  restarts_ = info.restarts;
  attempts_ = info.attempts;
  cur_beta_ = info.beta;
  if (info.have_best) record(info.best, best_state);
  device_check(qmc_walkers_pa_end(st.w), "qmc_walkers_pa_end");
  int32_t R = 0;
  qmc_walkers_count(st.w, &R);
  const int64_t N = model_->n_vars;
  std::vector<double> lowest(R);
  std::vector<uint8_t> ls((size_t)R * N);
  std::vector<uint64_t> acc(R);
  device_check(qmc_walkers_read(st.w, nullptr, lowest.data(), nullptr, acc.data(), nullptr, ls.data()),
               "qmc_walkers_read");
  device_check(qmc_walkers_energy(st.w, 1, lowest.data()), "qmc_walkers_energy");
  accepted_ = (int64_t)std::accumulate(acc.begin(), acc.end(), (uint64_t)0);
  // copy_solutions(Population): the global lowest first when no citizen holds it
  const bool had = have_lowest_;
  const double g = lowest_cost_;
  const std::vector<uint8_t> gs = lowest_state_;
  populate(lowest, ls, R);
  if (had && g < lowest_costs_.front()) {
    lowest_costs_.insert(lowest_costs_.begin(), g);
    lowest_states_.insert(lowest_states_.begin(), gs);
    if ((int)lowest_costs_.size() > number_of_solutions_) {
      lowest_costs_.pop_back();
      lowest_states_.pop_back();
    }
    lowest_cost_ = g;
    lowest_state_ = gs;
  }
  set_output("restarts", Json(restarts_));
  observe_run();
}

// The observations of every step (population_annealing.h:262-470) from the
// device's per-step records, in the reference's order: epoch, cost,
// population, sigma, tau and zeta where defined, and after a resampling the
// stepping size, family count, family sizes and rho; a restart restarts the
// observer's windows at its step.
void PopulationAnnealing::observe_run() {
  if (observer_.empty() || model_->empty()) return;
  PaState& st = pa_state(impl_);
  std::vector<double> log((size_t)steps_done_ * QMC_PA_LOG_FIELDS);
  device_check(qmc_walkers_pa_log(st.w, (int32_t)steps_done_, log.data()), "qmc_walkers_pa_log");
  observe_steps(steps_done_, [&](int64_t s) {
    const double* r = log.data() + s * QMC_PA_LOG_FIELDS;
    observer_.observe("epoch", r[0]);
    observer_.observe("cost", r[1]);
    observer_.observe("population", r[2]);
    observer_.observe("sigma", r[3]);
    if (r[4] == r[4]) {
      observer_.observe("tau", r[4]);
      observer_.observe("zeta", r[5]);
    }
    if (r[12] > 0) {
      observer_.observe("delta_beta", r[6]);
      observer_.observe("families", r[7]);
      observer_.observe_many("family_size", r[8], (int64_t)r[7], r[9]);
      observer_.observe("rho_t", r[10]);
    }
    if (r[11] > 0) observer_.restart(s);
  });
}

}  // namespace qiotk
