// Parameter-free simulated annealing and parallel tempering over fused
// device batches of worker runs (solver/parameter_free_solver.h:11-257,
// parameter_free_linear_solver.h:8-30, sa_parameter_free.h:18-175,
// pt_parameter_free.h:18-190, strategy/linear_search.cc:13-122).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <numeric>
#include <sstream>

#include "errors.h"
#include "solver.h"

namespace qiotk {

namespace {

const Json& params_of(const Json& doc) {
  if (doc.is_object() && doc.has("params")) return doc.at("params");
  return doc;
}

double wall_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Worker runs draw from their own counter streams: run k of a solve with
// seed s takes the seed splitmix64(s << 32 | k) >> 32.
uint32_t run_seed(uint32_t seed, int64_t k) {
  uint64_t z = ((uint64_t)seed << 32 | (uint64_t)(uint32_t)k) + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return (uint32_t)((z ^ (z >> 31)) >> 32);
}

template <class T>
[[noreturn]] void bad(const char* name, const char* rule, T bound, T found) {
  std::ostringstream o;
  o << "parameter `" << name << "`: must be " << rule << " " << bound << ", found " << found;
  throw Error(kValueError, o.str());
}

// Runs every worker's whole plan: those that fuse in one batch (one device
// launch per kernel variant), the rest one by one.  Returns each worker's
// share of the batch's device time (by replicas).
std::vector<double> run_workers(qmc_ctx ctx, const Model& m, std::vector<std::unique_ptr<Solver>>& ws) {
  std::vector<double> share(ws.size(), 0.0);
  qmc_batch b = nullptr;
  device_check(qmc_batch_create(ctx, &b), "qmc_batch_create");
  std::unique_ptr<qmc_batch_s, void (*)(qmc_batch)> hold(b, qmc_batch_destroy);
  std::vector<int> fused(ws.size(), 0);
  bool any = false;
  for (size_t k = 0; k < ws.size(); ++k) {
    ws[k]->init_with_retry();
    fused[k] = ws[k]->batch_plan(b) ? 1 : 0;
    any = any || (fused[k] && ws[k]->batch_replicas() > 0);
  }
  if (any) {
    device_check(qmc_batch_run(b, 1 << 30), "qmc_batch_run");
    double ms = 0;
    device_check(qmc_batch_wait(b, &ms), "qmc_batch_wait");
    int64_t tot_r = 0;
    for (auto& w : ws) tot_r += w->batch_replicas();
    std::vector<double> low(tot_r);
    std::vector<uint64_t> acc(tot_r);
    std::vector<uint8_t> st((size_t)tot_r * m.n_vars);
    device_check(qmc_batch_read(b, low.data(), acc.data(), st.data()), "qmc_batch_read");
    int64_t ro = 0;
    for (size_t k = 0; k < ws.size(); ++k) {
      const int64_t r = ws[k]->batch_replicas();
      if (!fused[k] || !r) continue;
      ws[k]->finalize_from(low.data() + ro, acc.data() + ro, st.data() + (size_t)ro * m.n_vars);
      share[k] = ms * (double)r / (double)tot_r;
      ro += r;
    }
  }
  for (size_t k = 0; k < ws.size(); ++k)
    if (!fused[k]) {
      ws[k]->run();
      ws[k]->finalize();
    }
  return share;
}

}  // namespace

std::string ParameterFree::target() const {
  return tempering_ ? "paralleltempering-parameterfree.qiotoolkit" : "simulatedannealing-parameterfree.qiotoolkit";
}

void ParameterFree::configure(const Json& doc) {
  // the worker's own configure comes first and checks its parameters
  // (sa_parameter_free.h:33, pt_parameter_free.h:43)
  {
    std::unique_ptr<Solver> w = create_solver(tempering_ ? "paralleltempering.qiotoolkit"
                                                         : "simulatedannealing.qiotoolkit");
    w->set_model(model_, dmodel_, ctx_);
    w->configure(doc);
  }
  Solver::configure(doc);
  const Json& p = params_of(doc);
  base_ = Json::object();
  // what a worker run takes from the request; its schedule, sweeps,
  // restarts, seed and limits come from the search
  for (const char* k : {"threads", "lanes", "positions_per_lane", "sweep_order", "replicas_per_group",
                        "number_of_solutions", "eval_limit", "enabled_features", "disabled_features",
                        "sweeps_per_replica"})
    if (p.has(k)) base_[k] = p.at(k);
  auto prob = [&](const char* name, double& v) {
    if (!p.has(name)) return;
    v = p.at(name).as_double();
    if (!(v > 0.0)) bad(name, "greater than", 0.0, v);
    if (!(v <= 1.0)) bad(name, "less-equal", 1.0, v);
  };
  prob("max_prob", max_prob_);
  prob("min_prob", min_prob_);
  prob("min_cost_diff_adjustment_factor", adjustment_);
  if (p.has("initial_sweeps_sample_points")) {
    const Json& a = p.at("initial_sweeps_sample_points");
    if (!a.is_array()) throw Error(kInvalidTypes, "parameter `initial_sweeps_sample_points` must be an array");
    points_.clear();
    for (const Json& e : a.items()) points_.push_back((int)e.as_int());
    if (points_.size() <= 2) {
      std::ostringstream o;
      o << "parameter `initial_sweeps_sample_points`: size must be greater than 2, found " << points_.size();
      throw Error(kValueError, o.str());
    }
    // a point of 0 would divide the restart count by zero
    for (int v : points_)
      if (v <= 0) bad("initial_sweeps_sample_points", "greater than", 0, v);
  }
  const char* nk = tempering_ ? "min_num_betas" : "min_num_restarts";
  // the reference defaults to its thread count; here SA takes the default
  // restart count and PT the longest ladder the device runs (32)
  min_num_ = threads_ > 0 ? threads_ : tempering_ ? 32 : 64;
  if (p.has(nk)) {
    min_num_ = p.at(nk).as_int();
    if (min_num_ <= 0) bad<int64_t>(nk, "greater than", 0, min_num_);
  }
  if (!model_->empty()) {
    max_diff_ = model_->estimate_max_cost_diff();
    min_diff_ = model_->estimate_min_cost_diff();
    std::ostringstream o;
    if (max_diff_ <= 0) o << "max cost difference should be > 0, found :" << max_diff_;
    else if (min_diff_ <= 0) o << "min cost difference should be > 0, found :" << min_diff_;
    else if (min_diff_ > max_diff_)
      o << "min cost difference should be <= max cost difference, found :" << min_diff_ << " > " << max_diff_;
    if (!o.str().empty()) throw Error(kInconsistentValue, o.str());
    beta_start_ = -std::log(max_prob_) / max_diff_;
    beta_stop_ = -std::log(min_prob_) / (min_diff_ + adjustment_ * (max_diff_ - min_diff_));
  }
  // a time limit must be set (parameter_free_solver.h:56-62)
  if (!time_limit_) time_limit_ = 5.0;
  step_limit_ = 0;  // the search stops on time or on no improvement
  out_params_ = output_for({(double)points_[0], (double)points_.back()});
  out_params_["seed"] = Json((int64_t)seed_);
}

// PTParameterFree::determine_all_betas_ (pt_parameter_free.h:150-172):
// geometric runs from beta_start to beta_stop of 4, 9, 14, ... values,
// concatenated until min_num_betas values.
std::vector<double> ParameterFree::betas() const {
  std::vector<double> b;
  Schedule g;
  Json j = Json::object();
  j["type"] = "geometric";
  j["initial"] = beta_start_;
  j["final"] = beta_stop_;
  for (int64_t size = 5;; size += 5) {
    const int64_t n = size - 1;
    j["start"] = Json((int64_t)0);
    j["stop"] = Json(n - 1);
    g.configure(j);
    for (int64_t i = 0; i < n; ++i) {
      b.push_back(g.get_value((double)i));
      if ((int64_t)b.size() >= min_num_) return b;
    }
  }
}

// SAParameterFree / PTParameterFree::update_parameters
// (sa_parameter_free.h:127-144, pt_parameter_free.h:117-137): sweeps =
// max(2, point); SA restarts = min_num_restarts * last point / point.
Json ParameterFree::output_for(const std::vector<double>& params) const {
  // a shrinking pass can take a small point of an unsorted list to 0
  const int64_t s = std::max<long>(1, std::lround(params[0])), back = std::lround(params[1]);
  Json o = Json::object();
  if (tempering_) {
    Json a = Json::array();
    for (double v : betas()) a.push_back(Json(v));
    o["all_betas"] = a;
  } else {
    o["beta_start"] = beta_start_;
    o["beta_stop"] = beta_stop_;
    o["restarts"] = Json(std::max<int64_t>(1, min_num_ * back / s));
  }
  o["sweeps"] = Json(std::max<int64_t>(s, 2));
  return o;
}

Json ParameterFree::worker_params(const std::vector<double>& params, uint32_t seed) const {
  Json p = output_for(params);
  for (const char* k : {"threads", "lanes", "positions_per_lane", "sweep_order", "replicas_per_group",
                        "number_of_solutions", "eval_limit", "enabled_features", "disabled_features",
                        "sweeps_per_replica"})
    if (base_.has(k)) p[k] = base_.at(k);
  p["seed"] = Json((int64_t)seed);
  Json doc = Json::object();
  doc["params"] = p;
  return doc;
}

std::unique_ptr<Solver> ParameterFree::worker(const std::vector<double>& params, uint32_t seed) const {
  std::unique_ptr<Solver> w =
      create_solver(tempering_ ? "paralleltempering.qiotoolkit" : "simulatedannealing.qiotoolkit");
  w->set_model(model_, dmodel_, ctx_);
  w->configure(worker_params(params, seed));
  return w;
}

// estimate_execution_cost (sa_parameter_free.h:146-152,
// pt_parameter_free.h:139-144): sweeps x replicas per thread
double ParameterFree::execution_cost(const std::vector<double>& params) const {
  const Json o = output_for(params);
  const double reps = tempering_ ? (double)o.at("all_betas").size() : (double)o.at("restarts").as_int();
  return (double)o.at("sweeps").as_int() * reps / (double)(threads_ > 0 ? threads_ : 64);
}

void ParameterFree::run() {
  if (model_->empty()) return;
  const double t0 = wall_s();
  // LinearSearchOpt state (strategy/linear_search.h:20-49)
  std::vector<int> points = points_;
  const int initial_min = points[0];
  std::vector<double> objective(points.size(), std::numeric_limits<double>::max());
  int cur = -1;
  int stable = 0;
  double exec_best = 0.0;
  double rate = 0.0;  // flip attempts per second of the last pass
  bool stop = false;
  have_lowest_ = false;
  while (!stop) {
    // the time check before a run (parameter_free_solver.h:101-108), here
    // before each pass
    if (have_lowest_ && wall_s() - t0 >= *time_limit_) {
      exit_reason_ = "Stop due to time limit: " + std::to_string(*time_limit_);
      break;
    }
    // LinearSearchOpt::fit_sample_points at the start of every pass after
    // the first (linear_search.cc:26-51)
    if (cur + 1 >= (int)points.size()) {
      const int lowest = (int)(std::min_element(objective.begin(), objective.end()) - objective.begin());
      if (lowest == 0) {
        if (points[0] > initial_min)
          for (int& v : points) v = (int)((double)v / change_factor_);
      } else if (lowest == (int)points.size() - 1) {
        for (int& v : points) v = (int)((double)v * change_factor_);
      }
      objective.assign(points.size(), std::numeric_limits<double>::max());
      cur = -1;
    }
    // the rest of the pass: parameters {point, last point} (linear_search.cc:53-93)
    std::vector<std::vector<double>> plan;
    for (int i = cur + 1; i < (int)points.size(); ++i) plan.push_back({(double)points[i], (double)points.back()});
    // The reference hands each run the time left as its own limit and cuts
    // it there; fused runs are not cut, so a pass that would not end in the
    // time left, at the rate the last pass ran, is not started.
    double planned = 0.0;
    for (const auto& pr : plan) {
      const Json o = output_for(pr);
      const double reps = tempering_ ? (double)o.at("all_betas").size() * (base_.has("sweeps_per_replica")
                                                                               ? (double)base_.at("sweeps_per_replica").as_int()
                                                                               : 1.0)
                                     : (double)o.at("restarts").as_int();
      planned += (double)o.at("sweeps").as_int() * reps * (double)model_->n_vars;
    }
    if (have_lowest_ && rate > 0 && wall_s() - t0 + planned / rate > *time_limit_) {
      exit_reason_ = "Stop due to time limit: " + std::to_string(*time_limit_) +
                     " (the next pass would not end within it)";
      break;
    }
    std::vector<std::unique_ptr<Solver>> ws;
    for (size_t k = 0; k < plan.size(); ++k) ws.push_back(worker(plan[k], run_seed(seed_, runs_ + (int64_t)k)));
    const double tp = wall_s();
    const std::vector<double> share = run_workers(ctx_, *model_, ws);
    rate = planned / std::max(1e-6, wall_s() - tp);
    ++passes_;
    for (size_t k = 0; k < ws.size(); ++k) {
      Solver& w = *ws[k];
      attempts_ += w.attempts_;
      accepted_ += w.accepted_;
      timing_.device_ms += w.timing_.device_ms + share[k];
      timing_.launches += w.timing_.launches;
      if (stop) {
        ++discarded_;
        continue;
      }
      ++runs_;
      ++steps_done_;
      ++cur;
      // ParameterFreeSolver::make_step (parameter_free_solver.h:110-157)
      const double obj = w.have_lowest_ ? w.lowest_cost_ : std::numeric_limits<double>::max();
      objective[cur] = obj;
      const double ec = execution_cost(plan[k]);
      auto update_best = [&] {
        lowest_cost_ = obj;
        lowest_state_ = w.lowest_state_;
        have_lowest_ = w.have_lowest_;
        best_params_ = plan[k];
        exec_best = ec;
      };
      if (have_lowest_ && std::fabs(lowest_cost_ - obj) < 10e-12) {
        ++stable;
        if (ec < exec_best) update_best();
      }
      if (stable >= nonimprovement_limit_) {
        exit_reason_ = "Stop after " + std::to_string(stable) + " runs without a lower cost";
        stop = true;
      }
      if (!have_lowest_ || lowest_cost_ - obj > 10e-12) {
        stable = 0;
        update_best();
      }
      // the lowest cost without the constant terms, as SteppingSolver
      // compares it (stepping_solver.h:151-152, model_solver.h:119)
      if (cost_limit_ && have_lowest_ && lowest_cost_ <= *cost_limit_) {
        exit_reason_ = "Stop due to cost limit: " + std::to_string(*cost_limit_);
        stop = true;
      }
      last_ = std::move(ws[k]);
    }
  }
  training_ = Json::object();
  training_["model_training_ms"] = Json((int64_t)0);
  training_["total_training_ms"] = Json((int64_t)0);
  training_["runs"] = Json(runs_);
  training_["passes"] = Json(passes_);
  training_["discarded_runs"] = Json(discarded_);
  training_["search_ms"] = (wall_s() - t0) * 1e3;
}

// ParameterFreeSolver::finalize (parameter_free_solver.h:160-177): the best
// run's lowest state first, then the last run's first solutions, and the
// best run's parameters as the output parameters
void ParameterFree::finalize() {
  if (model_->empty() || !last_) return;
  lowest_costs_.assign(1, lowest_cost_);
  lowest_states_.assign(1, lowest_state_);
  const size_t n = last_->lowest_costs_.size();
  for (size_t i = 0; i + 1 < n; ++i) {
    lowest_costs_.push_back(last_->lowest_costs_[i]);
    lowest_states_.push_back(last_->lowest_states_[i]);
  }
  if (!best_params_.empty()) {
    out_params_ = output_for(best_params_);
    out_params_["seed"] = Json((int64_t)seed_);
  }
}

Json ParameterFree::describe() const {
  Json d = Solver::describe();
  d["beta_start"] = beta_start_;
  d["beta_stop"] = beta_stop_;
  d["max_cost_diff"] = max_diff_;
  d["min_cost_diff"] = min_diff_;
  Json pts = Json::array();
  for (int v : points_) pts.push_back(Json((int64_t)v));
  d["initial_sweeps_sample_points"] = pts;
  d[tempering_ ? "min_num_betas" : "min_num_restarts"] = Json(min_num_);
  if (tempering_) {
    Json a = Json::array();
    for (double v : betas()) a.push_back(Json(v));
    d["all_betas"] = a;
  }
  d["time_limit"] = *time_limit_;
  return d;
}

// get_solver_properties (parameter_free_solver.h:185-192): the last worker
// run's properties plus "training"
Json ParameterFree::get_result() const {
  Json r = Solver::get_result();
  if (!last_) return r;
  Json ext = r["benchmark"]["extensions"];
  Json wext = last_->get_result()["benchmark"]["extensions"];
  Json out = Json::array();
  for (const Json& e : ext.items()) {
    if (e.at("name").as_string() != "solver") {
      out.push_back(e);
      continue;
    }
    Json sol;
    for (const Json& we : wext.items())
      if (we.at("name").as_string() == "solver") sol = we.at("value");
    sol["training"] = training_;
    if (!exit_reason_.empty()) sol["search_exit_reason"] = exit_reason_;
    Json c = Json::object();
    c["name"] = "solver";
    c["value"] = sol;
    out.push_back(c);
  }
  r["benchmark"]["extensions"] = out;
  return r;
}

}  // namespace qiotk
