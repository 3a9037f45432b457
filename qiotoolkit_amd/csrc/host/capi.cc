// C ABI of the host library (include/qiotk.h): Runner-style solves over the
// device engine.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "../../../include/qiotk.h"
#include "model.h"
#include "solver.h"

using namespace qiotk;

struct qiotk_engine_s {
  qmc_ctx ctx = nullptr;
};

struct qiotk_model_s {
  qiotk_engine e = nullptr;
  Model model;
  std::unique_ptr<DeviceModel> dev;
};

namespace {

thread_local std::string g_err;

char* dup(const std::string& s) {
  char* p = (char*)std::malloc(s.size() + 1);
  std::memcpy(p, s.c_str(), s.size() + 1);
  return p;
}

int report(const Error& e, char** response) {
  g_err = e.what();
  if (response) {
    Json j = Json::object();
    Json err = Json::object();
    err["code"] = Json((int64_t)e.code());
    err["message"] = std::string(e.what());
    j["error"] = err;
    *response = dup(j.dump());
  }
  return e.code();
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Runner::get_run_output (app/runner.cc:513-631): init, timed run, finalize.
Json run_one(Solver& s) {
  const double t0 = now_ms();
  s.init();
  const double t1 = now_ms();
  s.run();
  const double t2 = now_ms();
  s.finalize();
  Json r = s.get_result();
  r["benchmark"]["execution_time_ms"] = t2 - t1;
  r["benchmark"]["preprocessing_ms"] = t1 - t0;
  r["benchmark"]["postprocessing_ms"] = now_ms() - t2;
  r["benchmark"]["threads"] = Json((int64_t)1);
  return r;
}

template <class F>
int guarded(char** response, F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    return report(e, response);
  } catch (const JsonError& e) {
    return report(Error(kValueError, e.what()), response);
  } catch (const std::exception& e) {
    return report(Error(kRuntimeError, e.what()), response);
  }
}

}  // namespace

extern "C" {

const char* qiotk_last_error(void) { return g_err.c_str(); }

void qiotk_free(char* p) { std::free(p); }

int qiotk_engine_create(int device, qiotk_engine* out) {
  return guarded(nullptr, [&] {
    auto* e = new qiotk_engine_s();
    const int rc = qmc_ctx_create(device, &e->ctx);
    if (rc != QMC_OK) {
      delete e;
      device_check(rc, "qmc_ctx_create");
    }
    *out = e;
    return 0;
  });
}

void qiotk_engine_destroy(qiotk_engine e) {
  if (!e) return;
  qmc_ctx_destroy(e->ctx);
  delete e;
}

int qiotk_model_create(qiotk_engine e, const char* input_json, qiotk_model* out) {
  return guarded(nullptr, [&] {
    auto m = std::make_unique<qiotk_model_s>();
    m->e = e;
    m->model = parse_model(Json::parse(input_json));
    if (!m->model.empty()) m->dev = std::make_unique<DeviceModel>(e->ctx, m->model);
    *out = m.release();
    return 0;
  });
}

int qiotk_model_create_terms(qiotk_engine e, int kind, int64_t n_terms, const int64_t* offsets,
                             const int32_t* ids, const double* coeffs, qiotk_model* out) {
  return guarded(nullptr, [&] {
    std::vector<Term> terms(n_terms);
    for (int64_t k = 0; k < n_terms; ++k) {
      terms[k].c = coeffs[k];
      terms[k].ids.assign(ids + offsets[k], ids + offsets[k + 1]);
    }
    auto m = std::make_unique<qiotk_model_s>();
    m->e = e;
    m->model = make_model(kind == 0 ? QMC_ISING : QMC_PUBO, terms);
    if (!m->model.empty()) m->dev = std::make_unique<DeviceModel>(e->ctx, m->model);
    *out = m.release();
    return 0;
  });
}

void qiotk_model_destroy(qiotk_model m) { delete m; }

int qiotk_solve_model(qiotk_model m, const char* target, const char* params_json, char** response) {
  return guarded(response, [&] {
    auto s = create_solver(target);
    s->set_model(&m->model, m->dev.get(), m->e->ctx);
    s->configure(Json::parse(params_json));
    Json r = run_one(*s);
    *response = dup(r.dump());
    return 0;
  });
}

int qiotk_solve(qiotk_engine e, const char* target, const char* params_json, const char* input_json,
                char** response) {
  return guarded(response, [&] {
    const double t0 = now_ms();
    qiotk_model m = nullptr;
    int rc = qiotk_model_create(e, input_json, &m);
    if (rc) throw Error(rc, g_err);
    std::unique_ptr<qiotk_model_s> hold(m);
    auto s = create_solver(target);
    s->set_model(&m->model, m->dev.get(), e->ctx);
    s->configure(Json::parse(params_json));
    const double t1 = now_ms();
    Json r = run_one(*s);
    r["benchmark"]["configure_ms"] = t1 - t0;
    r["benchmark"]["end2end_time_ms"] = now_ms() - t0;
    *response = dup(r.dump());
    return 0;
  });
}

int qiotk_solve_batch(qiotk_engine e, const char* target, int32_t n, const qiotk_model* models,
                      const char* const* params_json, char** responses, double* device_ms) {
  for (int32_t k = 0; k < n; ++k) responses[k] = nullptr;
  return guarded(nullptr, [&] {
    std::vector<std::unique_ptr<Solver>> solvers;
    qmc_batch b = nullptr;
    device_check(qmc_batch_create(e->ctx, &b), "qmc_batch_create");
    std::unique_ptr<qmc_batch_s, void (*)(qmc_batch)> hold(b, qmc_batch_destroy);
    std::vector<int> fused(n, 0);
    const double t0 = now_ms();
    for (int32_t k = 0; k < n; ++k) {
      auto s = create_solver(target);
      s->set_model(&models[k]->model, models[k]->dev.get(), e->ctx);
      s->configure(Json::parse(params_json[k]));
      s->init();
      fused[k] = s->batch_plan(b) ? 1 : 0;
      solvers.push_back(std::move(s));
    }
    const double t1 = now_ms();
    double total_ms = 0;
    // every member finishes within its planned sweeps; one launch runs them all
    device_check(qmc_batch_run(b, 1 << 30), "qmc_batch_run");
    double ms = 0;
    device_check(qmc_batch_wait(b, &ms), "qmc_batch_wait");
    total_ms += ms;
    const double t2 = now_ms();
    for (int32_t k = 0; k < n; ++k) {
      if (!fused[k]) solvers[k]->run();
      solvers[k]->finalize();
    }
    const double t3 = now_ms();
    for (int32_t k = 0; k < n; ++k) {
      Json r = solvers[k]->get_result();
      r["benchmark"]["fused"] = Json((bool)fused[k]);
      r["benchmark"]["batch_init_ms"] = t1 - t0;
      r["benchmark"]["batch_run_ms"] = t2 - t1;
      r["benchmark"]["batch_finalize_ms"] = t3 - t2;
      r["benchmark"]["batch_device_ms"] = ms;
      responses[k] = dup(r.dump());
    }
    if (device_ms) *device_ms = total_ms;
    return 0;
  });
}

int qiotk_parse_model(const char* input_json, char** response) {
  return guarded(response, [&] {
    Model m = parse_model(Json::parse(input_json));
    Json j = Json::object();
    j["kind"] = m.kind == QMC_ISING ? "ising" : "pubo";
    j["n_vars"] = Json((int64_t)m.n_vars);
    j["const_cost"] = m.const_cost;
    Json names = Json::array(), off = Json::array(), vars = Json::array(), cs = Json::array();
    for (int32_t n : m.names) names.push_back(Json((int64_t)n));
    for (int64_t o : m.term_off) off.push_back(Json(o));
    for (int32_t v : m.term_var) vars.push_back(Json((int64_t)v));
    for (double c : m.coeff) cs.push_back(Json(c));
    j["names"] = names;
    j["term_offsets"] = off;
    j["term_vars"] = vars;
    j["coeffs"] = cs;
    j["scale_factor"] = m.scale_factor();
    Json ic = Json::array();
    for (auto& kv : m.initial) {
      Json e = Json::array();
      e.push_back(Json((int64_t)kv.first));
      e.push_back(Json((int64_t)kv.second));
      ic.push_back(e);
    }
    j["initial"] = ic;
    *response = dup(j.dump());
    return 0;
  });
}

int qiotk_schedule_eval(const char* schedule_json, int32_t n, const double* x, double* out,
                        char** discretized) {
  return guarded(discretized, [&] {
    Schedule s;
    s.configure(Json::parse(schedule_json));
    for (int32_t i = 0; i < n; ++i) out[i] = s.get_value(x[i]);
    if (discretized) {
      Json j = Json::object();
      j["start"] = s.get_start();
      j["stop"] = s.get_stop();
      j["count"] = Json((int64_t)s.get_count());
      Json vals = Json::array();
      try {
        for (double v : s.get_discretized_values()) vals.push_back(Json(v));
      } catch (const Error&) {
      }
      j["values"] = vals;
      *discretized = dup(j.dump());
    }
    return 0;
  });
}

int qiotk_describe(const char* target, const char* params_json, char** response) {
  return guarded(response, [&] {
    auto s = create_solver(target);
    Model empty;
    s->set_model(&empty, nullptr, nullptr);
    s->configure(Json::parse(params_json));
    *response = dup(s->describe().dump());
    return 0;
  });
}

}  // extern "C"
