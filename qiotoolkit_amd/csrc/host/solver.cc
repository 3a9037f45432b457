#include "solver.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>


namespace qiotk {

namespace {

double wall_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

const Json& params_of(const Json& doc) {
  if (!doc.is_object() || !doc.has("params"))
    throw Error(kMissingInput, "Input field `params` is required.");
  return doc.at("params");
}

// Metropolis::set_beta / set_temperature (markov/metropolis.h:35-61)
double beta_from(double v, bool use_inverse) {
  if (use_inverse) return v <= 0 ? 0.0 : v;
  return v <= 0 ? std::numeric_limits<double>::infinity() : 1.0 / v;
}

}  // namespace

// Replicas per work-group when the caller does not pick one: at least one
// work-group per compute unit, up to 8 replicas sharing each neighbour byte.
int default_group(int64_t R, int compute_units) {
  int B = 1;
  while (B < 8 && R / (2 * B) >= compute_units) B *= 2;
  return (int)std::min<int64_t>(B, R);
}

void Solver::init_with_retry() {
  try {
    init();
  } catch (const Error& e) {
    if (e.code() != kMemoryLimited || memory_saving_) throw;
    release_device();
    set_memory_saving(true);
    init();
  }
}

void Solver::set_memory_saving(bool on) {
  memory_saving_ = on;
  if (on) set_output("memory_saving", Json(true));
  else out_params_.remove("memory_saving");
}

int Solver::saving_group(int B, int divide) const {
  if (!memory_saving_) return B;
  int b = std::min(B, 8);
  while (divide > 0 && b > 1 && divide % b) --b;
  return std::max(1, b);
}

int Solver::compute_units() const {
  int32_t cus = 256;
  if (ctx_) qmc_ctx_compute_units(ctx_, &cus);
  return cus;
}

void Solver::configure(const Json& doc) {
  const Json& p = params_of(doc);
  observer_.configure(doc);
  // Solver::configure (solver.h:156-218)
  if (p.has("threads")) threads_ = (int)p.at("threads").as_int();
  if (p.has("lanes")) lanes_ = (int)p.at("lanes").as_int();
  if (p.has("sweeps_per_launch")) {
    sweeps_per_launch_ = p.at("sweeps_per_launch").as_int();
    if (sweeps_per_launch_ < 0) throw Error(kValueError, "`sweeps_per_launch` must be >= 0");
  }
  if (p.has("positions_per_lane")) {
    positions_per_lane_ = (int)p.at("positions_per_lane").as_int();
    if (positions_per_lane_ < 0) throw Error(kValueError, "`positions_per_lane` must be >= 0");
  }
  // Solver::configure (solver.h:205-217): feature lists; disabling wins
  auto features = [&](const char* key) {
    std::set<int> out;
    if (!p.has(key)) return out;
    const Json& a = p.at(key);
    if (!a.is_array()) throw Error(kInvalidTypes, std::string("parameter `") + key + "` must be an array");
    for (size_t i = 0; i < a.size(); ++i) out.insert((int)a[i].as_int());
    return out;
  };
  enabled_features_ = features("enabled_features");
  for (int f : features("disabled_features")) enabled_features_.erase(f);
  set_memory_saving(feature_enabled(3));  // FEATURE_USE_MEMORY_SAVING (utils/config.h:58)
  if (p.has("sweep_order")) {
    const std::string o = p.at("sweep_order").as_string();
    if (o == "colored") sweep_order_ = 0;
    else if (o == "index") sweep_order_ = 1;
    else if (o == "sequential") sweep_order_ = 2;
    else throw Error(kValueError, "`sweep_order` must be 'colored', 'index' or 'sequential'");
  }
  if (p.has("time_limit")) time_limit_ = p.at("time_limit").as_double();
  else if (p.has("timeout")) time_limit_ = p.at("timeout").as_double();
  if (p.has("cost_limit")) cost_limit_ = p.at("cost_limit").as_double();
  if (p.has("eval_limit")) eval_limit_ = p.at("eval_limit").as_int();
  if (p.has("number_of_solutions")) {
    const int64_t k = p.at("number_of_solutions").as_int();
    if (k < 1 || k > 1000) throw Error(kValueError, "`number_of_solutions` must be in [1, 1000]");
    number_of_solutions_ = (int)k;
  }
  // SteppingSolver::configure (stepping_solver.h:43-81)
  if (p.has("step_limit")) step_limit_ = p.at("step_limit").as_int();
  else if (p.has("sweeps")) step_limit_ = p.at("sweeps").as_int();
  if (p.has("steps_per_tick")) {
    steps_per_tick_ = p.at("steps_per_tick").as_int();
    if (steps_per_tick_ <= 0) throw Error(kValueError, "`steps_per_tick` must be greater than 0");
  }
  if (p.has("seed")) {
    seed_ = (uint32_t)p.at("seed").as_int();
  } else {
    seed_ = (uint32_t)(std::chrono::duration_cast<std::chrono::microseconds>(
                           std::chrono::system_clock::now().time_since_epoch())
                           .count());
  }
  set_output("seed", Json((int64_t)seed_));
  set_output("sweeps", Json(step_limit_));
}

int64_t Solver::steps_to_run() const { return std::max(step_limit_, steps_per_tick_); }

Json Solver::describe() const {
  Json d = Json::object();
  d["target"] = target();
  d["steps"] = Json(steps_to_run());
  d["seed"] = Json((int64_t)seed_);
  return d;
}

static Json to_json(const std::vector<double>& v) {
  Json a = Json::array();
  for (double x : v) a.push_back(Json(x));
  return a;
}

void Solver::record(double cost, const std::vector<uint8_t>& state) {
  // ModelSolver::update_lowest_cost (model_solver.h:247-258)
  if (!have_lowest_ || cost < lowest_cost_) {
    have_lowest_ = true;
    lowest_cost_ = cost;
    lowest_state_ = state;
  }
}

void Solver::populate(const std::vector<double>& lowest, const std::vector<uint8_t>& states,
                      int64_t R) {
  populate(lowest.data(), states.data(), R);
}

void Solver::populate(const double* lowest, const uint8_t* states, int64_t R) {
  // ModelSolver::copy_solutions + populate_solutions (model_solver.h:261-308)
  const int64_t N = model_->n_vars;
  std::vector<int64_t> idx(R);
  std::iota(idx.begin(), idx.end(), 0);
  const int64_t k = std::min<int64_t>(number_of_solutions_, R);
  std::partial_sort(idx.begin(), idx.begin() + k, idx.end(),
                    [&](int64_t a, int64_t b) { return lowest[a] < lowest[b]; });
  lowest_costs_.clear();
  lowest_states_.clear();
  for (int64_t i = 0; i < k; ++i) {
    lowest_costs_.push_back(lowest[idx[i]]);
    lowest_states_.emplace_back(states + idx[i] * N, states + (idx[i] + 1) * N);
  }
  if (k > 0) record(lowest_costs_[0], lowest_states_[0]);
}

int Solver::lanes_for(int replicas_per_group) const {
  // explicit lane count, else the smallest power-of-two multiple of 64 that
  // covers the largest colour class in `positions_per_lane` rounds
  if (lanes_) return lanes_;
  if (!positions_per_lane_ || !dmodel_) return 0;
  const int G = (replicas_per_group + 7) / 8;
  const int cap = std::max(64, 512 / G / 64 * 64);
  int l = 64;
  while (l < cap && (int64_t)l * positions_per_lane_ < dmodel_->info(visit()).max_class) l *= 2;
  return std::min(l, cap);
}

void Solver::apply_lanes(qmc_walkers w, int replicas_per_group) const {
  const int l = lanes_for(replicas_per_group);
  if (l) device_check(qmc_walkers_set_lanes(w, l), "qmc_walkers_set_lanes");
}

bool Solver::serve_plan(qmc_batch b) {
  if (served_) return false;
  served_ = true;
  serving_ = true;
  serve_t0_ = now_s();
  return batch_plan(b);
}

double Solver::now_s() { return wall_s(); }

bool Solver::limits_reached(int64_t steps, double t0, bool has_low, double low) {
  if (cost_limit_ && has_low && low <= *cost_limit_) {
    exit_reason_ = "Stop due to cost limit: " + std::to_string(*cost_limit_) + ", current lowest cost:";
    return true;
  }
  if (time_limit_ && wall_s() - t0 >= *time_limit_) {
    exit_reason_ = "Stop due to time_limit.steps : " + std::to_string(steps) + ", time_limit:" +
                   std::to_string(*time_limit_);
    return true;
  }
  return false;
}

// Serving: the whole run is planned at once; with time or cost limits they
// are checked after every wait for the device, and a solve past one
// completes there (the server takes its walkers out of the batch) with the
// steps it ran.
bool Solver::serve_done() {
  qmc_walkers w = walkers();
  if (!w) return true;
  int64_t left = 0;
  device_check(qmc_walkers_remaining(w, &left), "qmc_walkers_remaining");
  if (left == 0 || !(time_limit_ || cost_limit_)) return left == 0;
  const int64_t planned = steps_done_;
  const int64_t done = planned - (left + sweeps_per_step_ - 1) / sweeps_per_step_;
  double low = 0.0;
  if (cost_limit_) {
    std::vector<double> l((size_t)batched_replicas_);
    device_check(qmc_walkers_read(w, nullptr, l.data(), nullptr, nullptr, nullptr, nullptr), "qmc_walkers_read");
    low = *std::min_element(l.begin(), l.end());
  }
  if (!limits_reached(done, serve_t0_, cost_limit_.has_value(), low)) return false;
  if (planned > 0) attempts_ = attempts_ / planned * done;
  steps_done_ = done;
  return true;
}

void Solver::finalize_from(const double* lowest_energy, const uint64_t* accepted, const uint8_t* states) {
  const int64_t R = batched_replicas_;
  if (R == 0) return;
  accepted_ = (int64_t)std::accumulate(accepted, accepted + R, (uint64_t)0);
  populate(lowest_energy, states, R);
  observe_run();
}

bool Solver::best(double* cost, int8_t* config) const {
  if (!have_lowest_ || model_->empty()) {
    if (cost) *cost = model_->const_cost;
    return false;
  }
  if (cost) *cost = lowest_cost_ + model_->const_cost;
  if (config) {
    const bool ising = model_->kind == QMC_ISING;
    for (int32_t i = 0; i < model_->n_vars; ++i)
      config[i] = (int8_t)(ising ? (lowest_state_[i] ? -1 : 1) : (lowest_state_[i] ? 0 : 1));
  }
  return true;
}

Json Solver::get_solutions() const {
  Json s = Json::object();
  s["version"] = "1.0";
  if (have_lowest_ && !model_->empty()) {
    s["cost"] = lowest_cost_ + model_->const_cost;
    s["configuration"] = model_->render(lowest_state_.data());
    Json arr = Json::array();
    for (size_t i = 0; i < lowest_costs_.size(); ++i) {
      Json e = Json::object();
      e["cost"] = lowest_costs_[i] + model_->const_cost;
      e["configuration"] = model_->render(lowest_states_[i].data());
      arr.push_back(e);
    }
    s["solutions"] = arr;
  } else {
    s["cost"] = model_->const_cost;
  }
  s["parameters"] = out_params_;
  return s;
}

Json Solver::get_result() const {
  Json b = Json::object();
  // Solver::get_benchmark (solver/solver.h:78-146) with the graph size
  // property of graph/properties.h:52-69 under the keys of utils/config.h
  b["cost_function_evaluation_count"] = Json((int64_t)0);
  b["cost_difference_evaluation_count"] = Json(attempts_);
  Json size = Json::object();
  size["nodes"] = Json((int64_t)model_->n_vars);
  size["edges"] = Json(model_->n_terms());
  size["locality"] = Json((int64_t)model_->max_locality);
  size["min_locality"] = Json((int64_t)model_->min_locality);
  size["avg_locality"] = model_->n_terms() ? (double)model_->total_locality / (double)model_->n_terms() : 0.0;
  size["sum_coefficient_degrees_total"] = Json((int64_t)model_->total_locality);
  size["accumulated_dependent_vars"] = Json((int64_t)model_->dependent_vars);
  size["max_coupling_magnitude"] = model_->max_coupling;
  size["min_coupling_magnitude"] = model_->min_coupling;
  b["num_variables"] = size["nodes"];
  b["num_terms"] = size["edges"];
  for (const char* k : {"locality", "min_locality", "avg_locality", "accumulated_dependent_vars",
                        "sum_coefficient_degrees_total", "max_coupling_magnitude", "min_coupling_magnitude"})
    b[k] = size[k];
  Json ext = Json::array();
  Json counters = Json::object();
  counters["accepted_transitions"] = Json(accepted_);
  counters["difference_evaluations"] = Json(attempts_);
  counters["function_evaluations"] = Json((int64_t)0);
  Json c1 = Json::object();
  c1["name"] = "counters";
  c1["value"] = counters;
  ext.push_back(c1);
  // SteppingSolver::get_solver_properties (stepping_solver.h:268-283,
  // solver.h:57-75); milestones are recorded only in the reference's
  // profiling builds, so a release build reports them empty
  Json sol = observer_.render();
  if (time_limit_) sol["time_limit"] = *time_limit_;
  if (eval_limit_) sol["eval_limit"] = Json(*eval_limit_);
  if (cost_limit_) sol["cost_limit"] = *cost_limit_;
  sol["cost_milestones"] = Json::array();
  sol["last_step"] = Json(steps_done_);
  if (step_limit_ > 0) sol["step_limit"] = Json(step_limit_);
  std::string reason = exit_reason_;
  if (reason.empty() && step_limit_ > 0 && steps_done_ >= step_limit_)
    reason = "Stop due to step limit: " + std::to_string(step_limit_) + ", steps:" + std::to_string(steps_done_);
  if (!reason.empty()) sol["exit_reason"] = reason;
  Json c2 = Json::object();
  c2["name"] = "solver";
  c2["value"] = sol;
  ext.push_back(c2);
  // ModelSolver::get_model_properties -> GraphModel::get_benchmark_properties
  Json graph = Json::object();
  graph["size"] = size;
  Json mprops = Json::object();
  mprops["graph"] = graph;
  Json cm = Json::object();
  cm["name"] = "model";
  cm["value"] = mprops;
  ext.push_back(cm);
  Json dev = Json::object();
  dev["model"] = model_->kind == QMC_PUBO ? "pubo" : "ising";  // Model::get_identifier
  dev["device_time_ms"] = timing_.device_ms;
  dev["launches"] = Json(timing_.launches);
  dev["attempts_per_s"] = timing_.device_ms > 0 ? (double)attempts_ / (timing_.device_ms * 1e-3) : 0.0;
  if (dmodel_) {
    dev["colour_classes"] = Json((int64_t)dmodel_->info(visit()).n_classes);
    dev["coupling"] = dmodel_->info(visit()).coupling == QMC_COUPLING_I32 ? "int32" : "fp32";
  }
  Json c3 = Json::object();
  c3["name"] = "device";
  c3["value"] = dev;
  ext.push_back(c3);
  b["extensions"] = ext;
  Json r = Json::object();
  r["benchmark"] = b;
  r["solutions"] = get_solutions();
  return r;
}

// ===========================================================================
// SimulatedAnnealing (simulated_annealing.h:45-209)
void SimulatedAnnealing::configure(const Json& doc) {
  Solver::configure(doc);
  const Json& p = params_of(doc);
  use_inverse_temperature_ = p.has("use_inverse_temperature") && p.at("use_inverse_temperature").as_bool();
  bool qio_format = false;
  if (p.has("schedule")) {
    schedule_.configure(p.at("schedule"));
    set_output("schedule", schedule_.render());
  } else {
    const double b0 = p.has("beta_start") ? p.at("beta_start").as_double() : 0.0;
    const double b1 = p.has("beta_stop") ? p.at("beta_stop").as_double() : 0.01;
    if (b0 < 0) throw Error(kValueError, "parameter `beta_start` must be greater or equal 0");
    if (!(b1 > b0)) throw Error(kValueError, "parameter `beta_stop` must be greater than `beta_start`");
    set_output("beta_start", b0);
    set_output("beta_stop", b1);
    use_inverse_temperature_ = true;
    schedule_.make_linear(b0, b1);
    qio_format = true;
  }
  (void)qio_format;
  if (schedule_.get_start() == 0 && schedule_.get_stop() == 1 && schedule_.get_count() == 1) {
    if (step_limit_ > 1) schedule_.set_stop((int)std::max<int64_t>(1, step_limit_ - 1));
    else throw Error(kValueError, "The number of sweeps must be greater than 1.");
  }
  restarts_ = threads_ > 0 ? threads_ : 64;
  if (p.has("restarts")) {
    const int64_t r = p.at("restarts").as_int();
    if (r <= 0) throw Error(kValueError, "parameter `restarts` must be greater than 0");
    restarts_ = (int)r;
  }
  set_output("restarts", Json((int64_t)restarts_));
  if (p.has("replicas_per_group")) group_ = (int)p.at("replicas_per_group").as_int();
}

SimulatedAnnealing::~SimulatedAnnealing() {
  if (w_) qmc_walkers_destroy(w_);
}

void SimulatedAnnealing::release_device() {
  if (w_) qmc_walkers_destroy(w_);
  w_ = nullptr;
}

void SimulatedAnnealing::init() {
  if (model_->empty()) return;
  const bool seq = sequential();
  const int B = saving_group(group_ > 0 ? group_ : seq ? std::min(restarts_, 32)
                                                    : default_group(restarts_, compute_units()));
  device_check(qmc_walkers_create(dmodel_->handle(visit()), restarts_, B, seed_, 0, 0, restarts_, &w_),
               "qmc_walkers_create");
  apply_lanes(w_, B);
  if (seq) device_check(qmc_walkers_set_sweep_order(w_, 1), "qmc_walkers_set_sweep_order");
  if (model_->has_initial) {
    std::vector<uint8_t> st((size_t)restarts_ * model_->n_vars, 0);
    for (int r = 0; r < restarts_; ++r)
      for (auto& kv : model_->initial) st[(size_t)r * model_->n_vars + kv.first] = kv.second;
    device_check(qmc_walkers_set_states(w_, st.data()), "qmc_walkers_set_states");
  } else {
    device_check(qmc_walkers_init(w_, 0), "qmc_walkers_init");
  }
  // Batches of thousands of solves usually share one schedule: its table is
  // computed once and shared (keyed by the schedule, its length and its
  // beta/temperature reading).
  static std::mutex mu;
  static std::unordered_map<std::string, std::shared_ptr<const std::vector<double>>> cache;
  const std::string key = schedule_.key() + "|" + std::to_string(steps_to_run()) + "|" +
                          (use_inverse_temperature_ ? "b" : "t");
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      betas_ = it->second;
      return;
    }
  }
  auto b = std::make_shared<const std::vector<double>>(schedule_betas());
  std::lock_guard<std::mutex> lk(mu);
  if (cache.size() >= 64) cache.clear();
  cache.emplace(key, b);
  betas_ = std::move(b);
}

std::vector<double> SimulatedAnnealing::schedule_betas() const {
  // SimulatedAnnealing::make_step: one schedule value per step, as beta or T
  const int64_t steps = steps_to_run();
  std::vector<double> b(steps);
  for (int64_t s = 0; s < steps; ++s)
    b[s] = beta_from(schedule_.get_value((double)s), use_inverse_temperature_);
  return b;
}

Json SimulatedAnnealing::describe() const {
  Json d = Solver::describe();
  d["restarts"] = Json((int64_t)restarts_);
  d["betas"] = to_json(schedule_betas());
  return d;
}

// evaluations after s steps: one cost evaluation per restart at init, one
// difference evaluation per variable and restart per step
// (simulated_annealing.h:65-100)
int64_t SimulatedAnnealing::planned_steps() {
  const int64_t R = restarts_, N = model_->n_vars;
  return eval_limited_steps([&](int64_t s) { return R + s * R * N; });
}

void SimulatedAnnealing::run() {
  if (model_->empty()) return;
  const int64_t steps = planned_steps();
  const int64_t N = model_->n_vars;
  const double t0 = wall_s();
  // Without time or cost limits all steps are one launch; otherwise ticks of
  // steps with the limits checked in between (stepping_solver.h:127-258).
  const bool ticking = time_limit_ || cost_limit_;
  int64_t chunk = ticking ? std::max<int64_t>(1, steps_per_tick_) : steps;
  int64_t s = 0;
  while (s < steps) {
    const int64_t n = std::min(chunk, steps - s);
    device_check(qmc_walkers_sweep(w_, s, (int32_t)n, QMC_BETA_SCHEDULE, betas_->data() + s, 0),
                 "qmc_walkers_sweep");
    double ms = 0;
    qmc_walkers_last_ms(w_, &ms);
    timing_.device_ms += ms;
    timing_.launches++;
    s += n;
    attempts_ += n * N * restarts_;
    if (!ticking) continue;
    if (cost_limit_) {
      std::vector<double> low(restarts_);
      device_check(qmc_walkers_read(w_, nullptr, low.data(), nullptr, nullptr, nullptr, nullptr), "read");
      if (*std::min_element(low.begin(), low.end()) <= *cost_limit_) {
        exit_reason_ = "Stop due to cost limit: " + std::to_string(*cost_limit_) + ", current lowest cost:";
        break;
      }
    }
    if (time_limit_ && wall_s() - t0 >= *time_limit_) {
      exit_reason_ = "Stop due to time_limit.steps : " + std::to_string(s) + ", time_limit:" +
                     std::to_string(*time_limit_);
      break;
    }
    chunk = std::min<int64_t>(chunk * 2, 1 << 14);
  }
  steps_done_ = s;
}

bool SimulatedAnnealing::batch_plan(qmc_batch b) {
  // limits need checks between launches: served solves get them between
  // rounds, fused batches cannot stop one member early
  if ((time_limit_ || cost_limit_) && !serving_) return false;
  if (model_->empty()) return true;
  const int64_t steps = planned_steps();
  device_check(qmc_walkers_plan(w_, 0, (int32_t)steps, betas_->data()), "qmc_walkers_plan");
  device_check(qmc_batch_add(b, w_), "qmc_batch_add");
  batched_replicas_ = restarts_;
  steps_done_ = steps;
  attempts_ = steps * model_->n_vars * restarts_;
  return true;
}

void SimulatedAnnealing::finalize() {
  if (model_->empty()) return;
  const int64_t N = model_->n_vars;
  lowest_.assign(restarts_, 0.0);
  lowest_states_flat_.assign((size_t)restarts_ * N, 0);
  std::vector<uint64_t> acc(restarts_);
  device_check(qmc_walkers_read(w_, nullptr, lowest_.data(), nullptr, acc.data(), nullptr,
                                lowest_states_flat_.data()),
               "qmc_walkers_read");
  // exact FP64 energy of every returned configuration
  device_check(qmc_walkers_energy(w_, 1, lowest_.data()), "qmc_walkers_energy");
  accepted_ = (int64_t)std::accumulate(acc.begin(), acc.end(), (uint64_t)0);
  populate(lowest_, lowest_states_flat_, restarts_);
  observe_run();
}

// ===========================================================================
// ParallelTempering (parallel_tempering.h:68-234)

// The step observations of parallel_tempering.h:141-188: "step", then per
// replica (label "replica", the ladder position) its temperature or beta,
// and -- from the device's per-step records of an observed run -- its
// acceptance rate, cost and direction label, then each attempted pair's
// swap under the lower position's label.
void ParallelTempering::observe_run() {
  if (observer_.empty()) return;
  std::vector<double> t = temperatures_.get_discretized_values();
  if (use_inverse_temperatures_) std::sort(t.begin(), t.end(), std::greater<double>());
  else std::sort(t.begin(), t.end(), std::less<double>());
  const char* name = use_inverse_temperatures_ ? "beta" : "T";
  const int L = (int)t.size();
  std::vector<double> log;
  if (observe_dynamics_ && w_) {
    log.resize((size_t)steps_done_ * L * 4);
    device_check(qmc_walkers_pt_log(w_, (int64_t)log.size(), log.data()), "qmc_walkers_pt_log");
  }
  const double evals = (double)model_->n_vars * sweeps_per_replica_;
  observe_steps(steps_done_, [&](int64_t s) {
    const double* rec = log.empty() ? nullptr : log.data() + (size_t)s * L * 4;
    for (int i = 0; i < L; ++i) {
      observer_.set_label("replica", (int64_t)i);
      observer_.observe(name, use_inverse_temperatures_ ? beta_from(t[i], true) : (t[i] > 0 ? t[i] : 0.0));
      if (rec) {
        observer_.observe("acc_rate", rec[4 * i + 1] / evals);
        observer_.observe("avg_cost", rec[4 * i + 0]);
        observer_.observe("avg_dir", rec[4 * i + 2]);
      }
      observer_.clear_label("replica");
    }
    if (rec)
      for (int i = (int)(s & 1); i < L - 1; i += 2) {
        observer_.set_label("replica", (int64_t)i);
        observer_.observe("swap_rate", rec[4 * i + 3]);
        observer_.clear_label("replica");
      }
  });
}
void ParallelTempering::configure(const Json& doc) {
  Solver::configure(doc);
  const Json& p = params_of(doc);
  if (sequential())
    throw Error(kNotImplemented, "parallel tempering takes `sweep_order` 'colored' or 'index'");
  observe_dynamics_ = false;
  for (const char* name : {"acc_rate", "avg_cost", "avg_dir", "swap_rate"})
    observe_dynamics_ = observe_dynamics_ || observer_.watching(name);
  if (p.has("sweeps_per_replica")) {
    sweeps_per_replica_ = (int)p.at("sweeps_per_replica").as_int();
    if (sweeps_per_replica_ <= 0) throw Error(kValueError, "`sweeps_per_replica` must be greater than 0");
  }
  if (p.has("temperatures")) {
    temperatures_.configure(p.at("temperatures"));
    use_inverse_temperatures_ = false;
    set_output("temperatures", temperatures_.render());
  } else if (p.has("all_betas")) {
    temperatures_.configure(p.at("all_betas"));
    use_inverse_temperatures_ = true;
    set_output("all_betas", temperatures_.render());
  } else {
    temperatures_.make_list({0.1, 0.5, 1., 2., 4.});
    use_inverse_temperatures_ = true;
    set_output("all_betas", temperatures_.render());
  }
  if (p.has("replicas_per_group")) group_ = (int)p.at("replicas_per_group").as_int();
}

ParallelTempering::~ParallelTempering() {
  if (w_) qmc_walkers_destroy(w_);
}

void ParallelTempering::release_device() {
  if (w_) qmc_walkers_destroy(w_);
  w_ = nullptr;
}

std::vector<double> ParallelTempering::ladder() const {
  // ParallelTempering::init (parallel_tempering.h:68-81): betas descending or
  // temperatures ascending, then one beta per replica
  std::vector<double> t = temperatures_.get_discretized_values();
  if (use_inverse_temperatures_) std::sort(t.begin(), t.end(), std::greater<double>());
  else std::sort(t.begin(), t.end(), std::less<double>());
  std::vector<double> b(t.size());
  for (size_t k = 0; k < t.size(); ++k) b[k] = beta_from(t[k], use_inverse_temperatures_);
  return b;
}

Json ParallelTempering::describe() const {
  Json d = Solver::describe();
  d["betas"] = to_json(ladder());
  d["sweeps_per_replica"] = Json((int64_t)sweeps_per_replica_);
  return d;
}

void ParallelTempering::init() {
  if (model_->empty()) return;
  betas_ = ladder();
  const int L = (int)betas_.size();
  if (L > 32) throw Error(kValueError, "at most 32 temperatures per ladder are supported");
  // Default: split every ladder over work-groups of two replicas, so one
  // ladder occupies L/2 compute units that exchange through the in-kernel
  // ladder barrier (the best split measured on the 16^3 spin glass).
  int B = group_ > 0 ? group_ : ((L >= 4 && L % 2 == 0) ? 2 : L);
  if (L % B != 0) B = L;
  B = saving_group(B, L);
  device_check(qmc_walkers_create(dmodel_->handle(visit()), L, B, seed_, 0, 0, L, &w_), "qmc_walkers_create");
  apply_lanes(w_, B);
  if (observe_dynamics_)
    device_check(qmc_walkers_pt_observe(w_, L, planned_steps()), "qmc_walkers_pt_observe");
  if (model_->has_initial) {
    std::vector<uint8_t> st((size_t)L * model_->n_vars, 0);
    for (int r = 0; r < L; ++r)
      for (auto& kv : model_->initial) st[(size_t)r * model_->n_vars + kv.first] = kv.second;
    device_check(qmc_walkers_set_states(w_, st.data()), "qmc_walkers_set_states");
  } else {
    device_check(qmc_walkers_init(w_, 0), "qmc_walkers_init");
  }
}

// evaluations after s steps: one per replica at init, then every replica's
// difference evaluations (parallel_tempering.h:100-152)
int64_t ParallelTempering::planned_steps() {
  const int64_t L = (int64_t)betas_.size(), N = model_->n_vars;
  return eval_limited_steps([&](int64_t s) { return L + s * L * N * sweeps_per_replica_; });
}

void ParallelTempering::run() {
  if (model_->empty()) return;
  const int64_t steps = planned_steps();
  const int L = (int)betas_.size();
  const int64_t N = model_->n_vars;
  const double t0 = wall_s();
  const bool ticking = time_limit_ || cost_limit_;
  int64_t chunk = ticking ? std::max<int64_t>(1, steps_per_tick_) : steps;
  int64_t s = 0;
  while (s < steps) {
    const int64_t n = std::min(chunk, steps - s);
    device_check(qmc_walkers_pt(w_, L, betas_.data(), s, (int32_t)n, sweeps_per_replica_), "qmc_walkers_pt");
    double ms = 0;
    qmc_walkers_last_ms(w_, &ms);
    timing_.device_ms += ms;
    timing_.launches++;
    s += n;
    attempts_ += n * sweeps_per_replica_ * N * L;
    if (!ticking) continue;
    if (cost_limit_) {
      std::vector<double> low(L);
      device_check(qmc_walkers_read(w_, nullptr, low.data(), nullptr, nullptr, nullptr, nullptr), "read");
      if (*std::min_element(low.begin(), low.end()) <= *cost_limit_) {
        exit_reason_ = "Stop due to cost limit: " + std::to_string(*cost_limit_) + ", current lowest cost:";
        break;
      }
    }
    if (time_limit_ && wall_s() - t0 >= *time_limit_) {
      exit_reason_ = "Stop due to time_limit.steps : " + std::to_string(s) + ", time_limit:" +
                     std::to_string(*time_limit_);
      break;
    }
    chunk = std::min<int64_t>(chunk * 2, 1 << 14);
  }
  steps_done_ = s;
}

bool ParallelTempering::batch_plan(qmc_batch b) {
  // limits need checks between launches: served solves get them between
  // rounds, fused batches cannot stop one member early
  if (((time_limit_ || cost_limit_) && !serving_) || observe_dynamics_) return false;
  sweeps_per_step_ = sweeps_per_replica_;
  if (model_->empty()) return true;
  const int L = (int)betas_.size();
  const int64_t steps = planned_steps();
  if (qmc_walkers_plan_pt(w_, L, betas_.data(), 0, (int32_t)steps, sweeps_per_replica_) != QMC_OK)
    return false;  // ladder split over work-groups: runs on its own
  device_check(qmc_batch_add(b, w_), "qmc_batch_add");
  batched_replicas_ = L;
  steps_done_ = steps;
  attempts_ = steps * sweeps_per_replica_ * model_->n_vars * L;
  return true;
}

void ParallelTempering::finalize() {
  if (model_->empty()) return;
  const int L = (int)betas_.size();
  const int64_t N = model_->n_vars;
  lowest_.assign(L, 0.0);
  lowest_states_flat_.assign((size_t)L * N, 0);
  std::vector<uint64_t> acc(L);
  device_check(qmc_walkers_read(w_, nullptr, lowest_.data(), nullptr, acc.data(), nullptr,
                                lowest_states_flat_.data()),
               "qmc_walkers_read");
  device_check(qmc_walkers_energy(w_, 1, lowest_.data()), "qmc_walkers_energy");
  accepted_ = (int64_t)std::accumulate(acc.begin(), acc.end(), (uint64_t)0);
  populate(lowest_, lowest_states_flat_, L);
  observe_run();
}

// ===========================================================================
std::unique_ptr<Solver> create_solver(const std::string& target) {
  // create_model_solver (app/runner.cc:133-192)
  if (target == "simulatedannealing.qiotoolkit") return std::make_unique<SimulatedAnnealing>();
  if (target == "paralleltempering.qiotoolkit") return std::make_unique<ParallelTempering>();
  if (target == "populationannealing.cpu" || target == "populationannealing.qiotoolkit")
    return std::make_unique<PopulationAnnealing>();
  if (target == "simulatedannealing-parameterfree.qiotoolkit") return std::make_unique<ParameterFree>(false);
  if (target == "paralleltempering-parameterfree.qiotoolkit") return std::make_unique<ParameterFree>(true);
  throw Error(kValueError, "Unknown solver: " + target);
}

}  // namespace qiotk
