// PopulationAnnealing (solver/population_annealing.h:102-616) over the device
// engine: each step sweeps every citizen on the device, the host computes the
// Boltzmann weights and Poisson offspring counts, and the device copies the
// resampled population (Population::resample, solver/population.h:171-218).
// Random numbers follow the counter contract: citizens' sweeps are keyed by
// their slot and the global sweep index; offspring counts by (citizen, step).
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>

#include "rng.h"
#include "solver.h"

namespace qiotk {

namespace {
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
  if (!(alpha_ > 1.0)) throw Error(kValueError, "parameter `alpha` must be greater than 1");
  target_population_ = threads_ > 0 ? threads_ : 64;
  if (p.has("population")) target_population_ = p.at("population").as_int();
  if (target_population_ <= 0) throw Error(kValueError, "parameter `population` must be greater than 0");
  if (p.has("constant_population")) constant_population_ = p.at("constant_population").as_bool();
  if (constant_population_)
    throw Error(kNotImplemented, "constant_population resampling is not available on this engine");
  if (p.has("resampling_strategy")) strategy_ = p.at("resampling_strategy").as_string();
  if (strategy_ != "linear_schedule" && strategy_ != "geometric_schedule")
    throw Error(kNotImplemented, "resampling_strategy `" + strategy_ + "` is not available on this engine");
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
  if (p.has("sweeps_per_replica")) sweeps_per_replica_ = (int)p.at("sweeps_per_replica").as_int();
  if (sweeps_per_replica_ <= 0) throw Error(kValueError, "parameter `sweeps_per_replica` must be greater than 0");
  set_output("population", Json(target_population_));
  set_output("alpha", alpha_);
  set_output("sweeps_per_replica", Json((int64_t)sweeps_per_replica_));
}

PopulationAnnealing::~PopulationAnnealing() = default;

void PopulationAnnealing::init() {
  if (model_->empty()) return;
  PaState& st = pa_state(impl_);
  // Poisson resampling keeps the population near the target; room for
  // fluctuations is reserved up front.
  st.cap = (int32_t)std::max<int64_t>(4 * target_population_, target_population_ + 256);
  device_check(qmc_walkers_create(dmodel_->handle(), st.cap, 8, seed_, 0, 0, st.cap, &st.w),
               "qmc_walkers_create");
  std::vector<int32_t> ident(target_population_);
  std::iota(ident.begin(), ident.end(), 0);
  device_check(qmc_walkers_resample(st.w, ident.data(), (int32_t)target_population_), "resample");
  device_check(qmc_walkers_init(st.w, 0), "qmc_walkers_init");
  std::vector<double> b(target_population_, beta_start_);
  device_check(qmc_walkers_set_betas(st.w, b.data()), "qmc_walkers_set_betas");
  family_.resize(target_population_);
  std::iota(family_.begin(), family_.end(), (size_t)0);
  cur_beta_ = beta_start_;
}

void PopulationAnnealing::run() {
  if (model_->empty()) return;
  PaState& st = pa_state(impl_);
  const int64_t steps = steps_to_run();
  while (step_ < steps) {
    device_check(qmc_walkers_sweep(st.w, sweep_, sweeps_per_replica_, QMC_BETA_PER_REPLICA, nullptr, 0),
                 "qmc_walkers_sweep");
    double ms = 0;
    qmc_walkers_last_ms(st.w, &ms);
    timing_.device_ms += ms;
    timing_.launches++;
    finish_step();
  }
}

bool PopulationAnnealing::serve_plan(qmc_batch b) {
  if (model_->empty() || step_ >= steps_to_run()) return false;
  PaState& st = pa_state(impl_);
  device_check(qmc_walkers_plan(st.w, sweep_, sweeps_per_replica_, nullptr), "qmc_walkers_plan");
  if (!in_batch_) {
    device_check(qmc_batch_add(b, st.w), "qmc_batch_add");
    in_batch_ = true;
  }
  return true;
}

bool PopulationAnnealing::serve_after() {
  if (model_->empty()) return true;
  finish_step();
  return step_ >= steps_to_run();
}

qmc_walkers PopulationAnnealing::walkers() const {
  return impl_ ? static_cast<PaState*>(impl_.get())->w : nullptr;
}

// The host part of one step after its sweeps (population_annealing.h:262-488):
// lowest-cost update, weights, Poisson offspring, resampling and restarts.
void PopulationAnnealing::finish_step() {
  PaState& st = pa_state(impl_);
  const int64_t step = step_;
  const int64_t N = model_->n_vars;
  const double scale = model_->scale_factor();
  int32_t R = 0;
  qmc_walkers_count(st.w, &R);
  const double Rd = (double)R;
  sweep_ += sweeps_per_replica_;
  attempts_ += (int64_t)R * sweeps_per_replica_ * N;
  std::vector<double> after(R, 0.0), lowest(R, 0.0);
  device_check(qmc_walkers_read(st.w, after.data(), lowest.data(), nullptr, nullptr, nullptr, nullptr),
               "qmc_walkers_read");
  // update_lowest_cost over the citizens (:301-308)
  const int64_t kbest = std::min_element(lowest.begin(), lowest.end()) - lowest.begin();
  if (!have_lowest_ || lowest[kbest] < lowest_cost_) {
    std::vector<uint8_t> ls((size_t)N);
    device_check(qmc_walkers_read_replica(st.w, (int32_t)kbest, nullptr, ls.data()), "qmc_walkers_read_replica");
    record(lowest[kbest], ls);
  }
  double avg_after = 0, min_cost = std::numeric_limits<double>::max();
  for (double c : after) {
    avg_after += c;
    min_cost = c < min_cost ? c : min_cost;
  }
  avg_after /= Rd;
  double var_after = 0;
  for (double c : after) var_after += (c - avg_after) * (c - avg_after);
  var_after /= Rd;
  for (double& c : after) c -= min_cost;
  const double delta_beta = beta_.get_value((double)(step + 1 - base_step_)) -
                            beta_.get_value((double)(step - base_step_));
  double rho_t = 1.0;
  if (delta_beta > 0) {
    // resample (:490-616), Poisson branch
    std::vector<double> weight(R);
    double total = 0;
    for (int32_t i = 0; i < R; ++i) {
      weight[i] = std::exp(-delta_beta * after[i] * scale);
      total += weight[i];
    }
    const double RRQ = (double)target_population_ / total;
    std::vector<uint32_t> count(R, 1);
    int64_t current = R;
    for (int32_t i = 0; i < R; ++i) {
      const double mean = RRQ * weight[i];
      uint32_t nc = 0;
      if (mean > 0.0) {
        nc = poisson_draw(mean, [&](uint32_t k) {
          Words4 w = philox4x32_10((uint32_t)i, (uint32_t)step, k, DOM_POISSON, seed_, 0u);
          return canonical2(w.w[0], w.w[1]);
        });
      }
      if (nc == 0) {
        current -= count[i];
        count[i] = 0;
      } else if (nc > 1) {
        current += nc - 1;
        count[i] += nc - 1;
      }
    }
    if (current == 0) {  // keep the population alive
      Words4 w = philox4x32_10(0u, (uint32_t)step, 0u, DOM_PICK, seed_, 0u);
      const int32_t i = (int32_t)std::floor(canonical2(w.w[0], w.w[1]) * Rd);
      count[i] += 1;
      current = 1;
    }
    if (current > st.cap)
      throw Error(kMemoryLimited, "population grew beyond the reserved capacity");
    std::vector<int32_t> src;
    std::vector<size_t> fam;
    src.reserve(current);
    for (int32_t i = 0; i < R; ++i) {
      for (uint32_t k = 0; k < count[i]; ++k) {
        src.push_back(i);
        fam.push_back(family_[i]);
      }
    }
    device_check(qmc_walkers_resample(st.w, src.data(), (int32_t)src.size()), "qmc_walkers_resample");
    family_.swap(fam);
    std::map<size_t, size_t> census;
    for (size_t f : family_) census[f]++;
    rho_t = 0;
    for (auto& kv : census) rho_t += (double)kv.second * (double)kv.second;
    rho_t /= Rd;
  }
  if (Rd / rho_t < alpha_ || var_after == 0) {
    // restart (:461-475): fresh random citizens; existing ones keep their
    // beta until the next step sets it, as init_population does
    cur_beta_ = beta_start_;
    base_step_ = step;
    int32_t Rnow = 0;
    qmc_walkers_count(st.w, &Rnow);
    if (Rnow < target_population_) {
      std::vector<double> b(Rnow);
      device_check(qmc_walkers_read(st.w, nullptr, nullptr, b.data(), nullptr, nullptr, nullptr), "read");
      std::vector<int32_t> src(target_population_);
      for (int64_t i = 0; i < target_population_; ++i) src[i] = (int32_t)std::min<int64_t>(i, Rnow - 1);
      device_check(qmc_walkers_resample(st.w, src.data(), (int32_t)target_population_), "resample");
      b.resize(target_population_, beta_start_);
      device_check(qmc_walkers_set_betas(st.w, b.data()), "qmc_walkers_set_betas");
      Rnow = (int32_t)target_population_;
    }
    device_check(qmc_walkers_init(st.w, ++epoch_), "qmc_walkers_init");
    family_.resize(Rnow);
    for (int64_t i = 0; i < Rnow; ++i)
      if (i < target_population_) family_[i] = (size_t)i;
    restarts_++;
  } else {
    cur_beta_ += delta_beta;
    int32_t Rnow = 0;
    qmc_walkers_count(st.w, &Rnow);
    std::vector<double> b(Rnow, cur_beta_);
    device_check(qmc_walkers_set_betas(st.w, b.data()), "qmc_walkers_set_betas");
  }
  step_ = step + 1;
  steps_done_ = step_;
}

void PopulationAnnealing::finalize() {
  if (model_->empty()) return;
  PaState& st = pa_state(impl_);
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
  Json ext = Json::object();
  set_output("restarts", Json(restarts_));
}

}  // namespace qiotk
