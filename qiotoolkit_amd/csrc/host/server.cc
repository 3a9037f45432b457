// Serving loop: solve requests arrive at any time, are admitted while their
// device footprint fits the HBM budget, and every admitted solve advances in
// shared launches, round by round, with its state resident on the device
// until it completes (continuous batching).
//
// The reference has no server; one Runner process solves one request
// (app/runner.cc:513-688).  Each request here is that solve -- the same
// targets, params and response document -- and its result does not depend on
// what else is in flight: RNG counters are keyed by the request's own seed,
// replica and sweep, and the host parts of a step (population annealing's
// resampling) use only the request's own state.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../err_buf.h"
#include "../../../include/qiotk.h"
#include "model.h"
#include "solver.h"

using namespace qiotk;

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

char* dup(const std::string& s) {
  char* p = (char*)std::malloc(s.size() + 1);
  std::memcpy(p, s.c_str(), s.size() + 1);
  return p;
}

thread_local qmc::ErrBuf g_err;

Json error_doc(int code, const std::string& msg) {
  Json j = Json::object();
  Json e = Json::object();
  e["code"] = Json((int64_t)code);
  e["message"] = msg;
  j["error"] = e;
  return j;
}

struct Request {
  int64_t id = 0;
  std::string target;
  Json params;
  Model model;
  double arrival = 0;  // ms on the server clock
};

struct Job {
  std::unique_ptr<Request> req;
  std::unique_ptr<DeviceModel> dev;
  std::unique_ptr<Solver> solver;
  int64_t footprint = 0;
  double admitted = 0;
  int64_t rounds = 0;
  bool in_batch = false;
};

struct Done {
  std::string response;
  int code = 0;
};

// Device bytes a solve holds while admitted: the problem's term graph and
// sweep layouts, the walkers' packed spins (current + lowest), per-replica
// arrays and the plain-state staging used by reads and resampling.
int64_t footprint(const Model& m, const std::string& target, int64_t cap) {
  if (m.empty()) return 0;
  const int64_t N = m.n_vars, M = m.n_terms();
  const int64_t inc = (int64_t)m.term_var.size();
  const int64_t problem = 48 * inc + 24 * M + 24 * N + (1 << 16);
  const int64_t staging = (target.rfind("populationannealing", 0) == 0 ? 2 : 1) * cap * N;
  // current + lowest states: a work-group of B replicas holds ceil(B/8)
  // bytes per position, so at most one byte per replica and position
  const int64_t spins = 2 * cap * (N + 4);
  return problem + spins + staging + 64 * cap;
}

}  // namespace

struct qiotk_server_s {
  qmc_ctx ctx = nullptr;
  qmc_batch batch = nullptr;
  int64_t budget = 64ll << 30;
  int max_active = 64;
  int sweeps_per_round = 10;
  // rounds enqueued per wait for the device: adapted so that about
  // round_target_ms of device work is in flight, at most max_rounds_per_wait
  double round_target_ms = 2.0;
  int max_rounds_per_wait = 16;
  int rounds_per_wait = 1;
  double t0 = 0;

  std::mutex mu;
  std::condition_variable cv_in, cv_out;
  std::deque<std::unique_ptr<Request>> incoming;
  std::map<int64_t, Done> done;
  int64_t next_id = 1;
  bool stop = false;

  // owned by the loop thread
  std::deque<std::unique_ptr<Request>> pending;
  std::vector<std::unique_ptr<Job>> active;
  int64_t used = 0;

  // statistics (under mu)
  std::vector<double> latencies, queue_ms;
  int64_t completed = 0, failed = 0, rounds = 0, attempts = 0;
  double device_ms = 0, first_arrival = -1, last_finish = 0, active_sum = 0;
  int64_t peak_used = 0, peak_active = 0, peak_pending = 0;

  std::thread loop;

  void finish(Job& j, Json result, int code) {
    const double t = now_ms() - t0;
    if (code == 0) {
      Json& b = result["benchmark"];
      Json srv = Json::object();
      srv["arrival_ms"] = j.req->arrival;
      srv["admitted_ms"] = j.admitted;
      srv["finished_ms"] = t;
      srv["latency_ms"] = t - j.req->arrival;
      srv["queue_ms"] = j.admitted - j.req->arrival;
      srv["rounds"] = Json(j.rounds);
      srv["device_bytes_reserved"] = Json(j.footprint);
      b["serving"] = srv;
    }
    std::lock_guard<std::mutex> lk(mu);
    done[j.req->id] = Done{result.dump(), code};
    if (code == 0) {
      completed++;
      latencies.push_back(t - j.req->arrival);
      queue_ms.push_back(j.admitted - j.req->arrival);
      if (j.solver) attempts += j.solver->attempts();
    } else {
      failed++;
    }
    last_finish = std::max(last_finish, t);
    cv_out.notify_all();
  }

  void fail_request(std::unique_ptr<Request> r, int code, const std::string& msg) {
    Job j;
    j.req = std::move(r);
    finish(j, error_doc(code, msg), code);
  }

  // Admit pending requests in arrival order while slots and memory allow.
  void admit() {
    while (!pending.empty() && (int)active.size() < max_active) {
      Request& r = *pending.front();
      int64_t fp = 0;
      bool saving = false;
      std::unique_ptr<Solver> s;
      try {
        s = create_solver(r.target);
        if (!s->servable())
          throw Error(kNotImplemented, "the serving loop does not run target " + r.target);
        s->configure(r.params);
        saving = s->memory_saving();
        fp = footprint(r.model, r.target, s->device_replicas());
        if (fp > budget && !saving) {
          // Runner::get_run_output's retry (app/runner.cc:518-537) at
          // admission: the memory-saving form, when that one fits
          s->set_memory_saving(true);
          const int64_t fp_saving = footprint(r.model, r.target, s->device_replicas());
          if (fp_saving <= budget) {
            fp = fp_saving;
            saving = true;
          }
        }
      } catch (const Error& e) {
        auto rq = std::move(pending.front());
        pending.pop_front();
        fail_request(std::move(rq), e.code(), e.what());
        continue;
      } catch (const std::exception& e) {
        auto rq = std::move(pending.front());
        pending.pop_front();
        fail_request(std::move(rq), kValueError, e.what());
        continue;
      }
      if (fp > budget) {
        auto rq = std::move(pending.front());
        pending.pop_front();
        fail_request(std::move(rq), kMemoryLimited, "request needs more device memory than the serving budget");
        continue;
      }
      if (used + fp > budget) break;  // wait for memory to free up
      auto job = std::make_unique<Job>();
      job->req = std::move(pending.front());
      pending.pop_front();
      job->footprint = fp;
      job->admitted = now_ms() - t0;
      try {
        if (!job->req->model.empty()) job->dev = std::make_unique<DeviceModel>(ctx, job->req->model);
        job->solver = create_solver(job->req->target);
        job->solver->set_model(&job->req->model, job->dev.get(), ctx);
        job->solver->configure(job->req->params);
        if (saving) job->solver->set_memory_saving(true);
        job->solver->init_with_retry();
      } catch (const Error& e) {
        finish(*job, error_doc(e.code(), e.what()), e.code());
        continue;
      } catch (const std::exception& e) {
        finish(*job, error_doc(kRuntimeError, e.what()), kRuntimeError);
        continue;
      }
      used += fp;
      active.push_back(std::move(job));
    }
    std::lock_guard<std::mutex> lk(mu);
    peak_used = std::max(peak_used, used);
    peak_active = std::max<int64_t>(peak_active, (int64_t)active.size());
  }

  void complete(std::unique_ptr<Job>& jp) {
    Job& j = *jp;
    qmc_walkers w = j.solver ? j.solver->walkers() : nullptr;
    if (w && j.in_batch) qmc_batch_remove(batch, w);
    try {
      j.solver->finalize();
      finish(j, j.solver->get_result(), 0);
    } catch (const Error& e) {
      finish(j, error_doc(e.code(), e.what()), e.code());
    } catch (const std::exception& e) {
      finish(j, error_doc(kRuntimeError, e.what()), kRuntimeError);
    }
    used -= j.footprint;
    jp.reset();
  }

  // A solve that failed mid-flight: out of the batch, its error as the result.
  void drop(std::unique_ptr<Job>& jp, int code, const std::string& msg) {
    Job& j = *jp;
    if (j.in_batch && j.solver && j.solver->walkers()) qmc_batch_remove(batch, j.solver->walkers());
    finish(j, error_doc(code, msg), code);
    used -= j.footprint;
    jp.reset();
  }

  void run_loop() {
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mu);
        if (active.empty() && pending.empty() && incoming.empty()) {
          if (stop) break;
          cv_in.wait_for(lk, std::chrono::milliseconds(20));
        }
        while (!incoming.empty()) {
          pending.push_back(std::move(incoming.front()));
          incoming.pop_front();
        }
        peak_pending = std::max<int64_t>(peak_pending, (int64_t)pending.size());
      }
      admit();
      if (active.empty()) continue;
      // Enqueue rounds -- plan every active solve, run the batch, enqueue
      // what follows each solve's round -- and wait for the device once.
      int rc = QMC_OK;
      int sub = 0;
      while (sub < rounds_per_wait && rc == QMC_OK) {
        for (auto& jp : active) {
          Job& j = *jp;
          try {
            const bool added = j.solver->serve_plan(batch);
            if (added) j.in_batch = true;
            if (!added && !j.in_batch) {
              // not expressible as rounds (time or cost limits): solve now
              j.solver->run();
              complete(jp);
            }
          } catch (const Error& e) {
            drop(jp, e.code(), e.what());
          } catch (const std::exception& e) {
            drop(jp, kRuntimeError, e.what());
          }
        }
        active.erase(std::remove(active.begin(), active.end(), nullptr), active.end());
        if (active.empty()) break;
        rc = qmc_batch_run(batch, sweeps_per_round);
        if (rc != QMC_OK) break;
        ++sub;
        for (auto& jp : active) {
          try {
            jp->solver->serve_advance();
          } catch (const Error& e) {
            drop(jp, e.code(), e.what());
          } catch (const std::exception& e) {
            drop(jp, kRuntimeError, e.what());
          }
        }
        active.erase(std::remove(active.begin(), active.end(), nullptr), active.end());
      }
      if (sub == 0 && rc == QMC_OK) continue;
      double ms = 0;
      if (rc == QMC_OK) rc = qmc_batch_wait(batch, &ms);
      {
        std::lock_guard<std::mutex> lk(mu);
        rounds += sub;
        device_ms += ms;
        active_sum += (double)active.size() * sub;
      }
      if (rc == QMC_OK && sub > 0 && ms > 0) {
        const double per = ms / sub;
        rounds_per_wait = std::max(1, std::min(max_rounds_per_wait, (int)(round_target_ms / per)));
      }
      for (auto& jp : active) {
        Job& j = *jp;
        j.rounds += sub;
        if (rc != QMC_OK) {
          drop(jp, rc, qmc_last_error());
          continue;
        }
        bool finished = false;
        try {
          finished = j.solver->serve_done();
        } catch (const Error& e) {
          drop(jp, e.code(), e.what());
          continue;
        } catch (const std::exception& e) {
          drop(jp, kRuntimeError, e.what());
          continue;
        }
        if (finished) complete(jp);
      }
      active.erase(std::remove(active.begin(), active.end(), nullptr), active.end());
    }
  }
};

extern "C" {

int qiotk_server_create(int device, const char* config_json, qiotk_server* out) {
  try {
    auto s = std::make_unique<qiotk_server_s>();
    if (config_json && *config_json) {
      Json c = Json::parse(config_json);
      if (c.has("hbm_budget_bytes")) s->budget = c.at("hbm_budget_bytes").as_int();
      if (c.has("max_active")) s->max_active = (int)c.at("max_active").as_int();
      if (c.has("sweeps_per_round")) s->sweeps_per_round = (int)c.at("sweeps_per_round").as_int();
      if (c.has("round_target_ms")) s->round_target_ms = c.at("round_target_ms").as_double();
      if (c.has("max_rounds_per_wait")) s->max_rounds_per_wait = (int)c.at("max_rounds_per_wait").as_int();
      if (s->budget <= 0 || s->max_active <= 0 || s->sweeps_per_round <= 0 || s->max_rounds_per_wait <= 0 ||
          !(s->round_target_ms > 0))
        throw Error(kValueError,
                    "hbm_budget_bytes, max_active, sweeps_per_round, round_target_ms and max_rounds_per_wait "
                    "must be positive");
    }
    device_check(qmc_ctx_create(device, &s->ctx), "qmc_ctx_create");
    device_check(qmc_batch_create(s->ctx, &s->batch), "qmc_batch_create");
    s->t0 = now_ms();
    qiotk_server_s* raw = s.get();
    s->loop = std::thread([raw] { raw->run_loop(); });
    *out = s.release();
    return 0;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code();
  } catch (const std::exception& e) {
    g_err = e.what();
    return kValueError;
  }
}

// arrival = when the call began, so host-side parsing counts toward latency
static int submit(qiotk_server s, const char* target, const char* params_json, Model&& m, double arrival,
                  int64_t* id) {
  auto r = std::make_unique<Request>();
  r->target = target;
  r->params = Json::parse(params_json);
  r->model = std::move(m);
  std::lock_guard<std::mutex> lk(s->mu);
  r->arrival = arrival;
  if (s->first_arrival < 0) s->first_arrival = r->arrival;
  r->id = s->next_id++;
  *id = r->id;
  s->incoming.push_back(std::move(r));
  s->cv_in.notify_one();
  return 0;
}

int qiotk_server_submit(qiotk_server s, const char* target, const char* params_json, int kind,
                        int64_t n_terms, const int64_t* offsets, const int32_t* ids, const double* coeffs,
                        int64_t* request_id) {
  const double arrival = now_ms() - s->t0;
  try {
    Model m = make_model_flat(kind == 0 ? QMC_ISING : QMC_PUBO, n_terms, offsets, ids, coeffs);
    return submit(s, target, params_json, std::move(m), arrival, request_id);
  } catch (const Error& e) {
    g_err = e.what();
    return e.code();
  } catch (const std::exception& e) {
    g_err = e.what();
    return kValueError;
  }
}

int qiotk_server_submit_json(qiotk_server s, const char* target, const char* params_json,
                             const char* input_json, int64_t* request_id) {
  const double arrival = now_ms() - s->t0;
  try {
    Model m = read_model_text(input_json);
    return submit(s, target, params_json, std::move(m), arrival, request_id);
  } catch (const Error& e) {
    g_err = e.what();
    return e.code();
  } catch (const std::exception& e) {
    g_err = e.what();
    return kValueError;
  }
}

int qiotk_server_wait(qiotk_server s, int64_t request_id, double timeout_ms, char** response) {
  *response = nullptr;
  std::unique_lock<std::mutex> lk(s->mu);
  auto ready = [&] { return s->done.count(request_id) > 0; };
  if (timeout_ms < 0) s->cv_out.wait(lk, ready);
  else if (!s->cv_out.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) return 0;
  Done d = std::move(s->done[request_id]);
  s->done.erase(request_id);
  *response = dup(d.response);
  return d.code;
}

static double pct(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

int qiotk_server_stats(qiotk_server s, char** response) {
  std::lock_guard<std::mutex> lk(s->mu);
  Json j = Json::object();
  j["completed"] = Json(s->completed);
  j["failed"] = Json(s->failed);
  j["rounds"] = Json(s->rounds);
  j["attempts"] = Json(s->attempts);
  j["device_ms"] = s->device_ms;
  j["latency_p50_ms"] = pct(s->latencies, 0.50);
  j["latency_p99_ms"] = pct(s->latencies, 0.99);
  j["latency_max_ms"] = s->latencies.empty() ? 0.0 : *std::max_element(s->latencies.begin(), s->latencies.end());
  j["queue_p50_ms"] = pct(s->queue_ms, 0.50);
  j["queue_p99_ms"] = pct(s->queue_ms, 0.99);
  const double window = s->first_arrival >= 0 ? s->last_finish - s->first_arrival : 0.0;
  j["window_ms"] = window;
  j["requests_per_s"] = window > 0 ? (double)s->completed / (window * 1e-3) : 0.0;
  j["attempts_per_s"] = window > 0 ? (double)s->attempts / (window * 1e-3) : 0.0;
  j["mean_active"] = s->rounds ? s->active_sum / (double)s->rounds : 0.0;
  j["peak_active"] = Json(s->peak_active);
  j["peak_pending"] = Json(s->peak_pending);
  j["peak_reserved_bytes"] = Json(s->peak_used);
  j["hbm_budget_bytes"] = Json(s->budget);
  j["sweeps_per_round"] = Json((int64_t)s->sweeps_per_round);
  j["rounds_per_wait"] = Json((int64_t)s->rounds_per_wait);
  *response = dup(j.dump());
  return 0;
}

void qiotk_server_destroy(qiotk_server s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_in.notify_all();
  }
  if (s->loop.joinable()) s->loop.join();
  s->active.clear();
  qmc_batch_destroy(s->batch);
  qmc_ctx_destroy(s->ctx);
  delete s;
}

const char* qiotk_server_last_error(void) { return g_err.c_str(); }

}  // extern "C"
