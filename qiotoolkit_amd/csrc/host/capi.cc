// C ABI of the host library (include/qiotk.h): Runner-style solves over the
// device engine.
#include <sys/resource.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <dirent.h>
#include <memory>
#include <string>

#include "../err_buf.h"
#include "../../../include/qiotk.h"
#include "model.h"
#include "observer.h"
#include "proto.h"
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

thread_local qmc::ErrBuf g_err;

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

// process CPU time (utils/timing.cc:29)
double cpu_ms() { return 1e3 * (double)std::clock() / CLOCKS_PER_SEC; }

// peak resident memory (utils/operating_system.cc:20-25)
int64_t max_rss_bytes() {
  struct rusage ru {};
  getrusage(RUSAGE_SELF, &ru);
  return (int64_t)ru.ru_maxrss * 1024;
}

// Bytes of an input file, or of every file of a protobuf problem folder.
int64_t path_bytes(const char* path) {
  struct stat st {};
  if (stat(path, &st) != 0) return 0;
  if (!S_ISDIR(st.st_mode)) return (int64_t)st.st_size;
  int64_t n = 0;
  if (DIR* d = opendir(path)) {
    while (dirent* e = readdir(d)) {
      struct stat fs {};
      const std::string f = std::string(path) + "/" + e->d_name;
      if (stat(f.c_str(), &fs) == 0 && S_ISREG(fs.st_mode)) n += (int64_t)fs.st_size;
    }
    closedir(d);
  }
  return n;
}

const double kStartupMs = now_ms();
const std::time_t kStartup = std::time(nullptr);

// utils/metadata.cc:142-161
Json build_properties() {
  Json b = Json::object();
  b["branch"] = std::string("main");
  b["commit_hash"] = std::string(QIOTK_COMMIT);
  b["compiler"] = std::string("g++ " __VERSION__);
  b["build_type"] = std::string("Release");
  b["qiotoolkit_version"] = std::string("qiotoolkit_amd (MI355X)");
  return b;
}
Json invocation_properties() {
  Json v = Json::object();
  char host[256] = {0};
  gethostname(host, sizeof host - 1);
  v["host"] = std::string(host);
  const char* user = std::getenv("USER");
  v["user"] = std::string(user ? user : "");
  char when[64] = {0};
  std::strftime(when, sizeof when, "%Y-%m-%d %H:%M:%S", std::localtime(&kStartup));
  v["datetime"] = std::string(when);
  char cwd[4096] = {0};
  v["directory"] = std::string(getcwd(cwd, sizeof cwd - 1) ? cwd : "");
  return v;
}

// Where and when a request started, and what it read.
struct Request {
  double t0_ms = now_ms();
  double cpu0_ms = cpu_ms();
  int64_t read_bytes = 0;
};

// Runner::get_run_output (app/runner.cc:513-631): init, timed run,
// finalize, and the benchmark block under the keys of utils/config.h.
Json run_one(Solver& s, const Request& rq) {
  const double t0 = now_ms(), c0 = cpu_ms();
  s.init_with_retry();
  const double t1 = now_ms();
  s.run();
  const double t2 = now_ms();
  s.finalize();
  Json r = s.get_result();
  Json& b = r["benchmark"];
  // the solver's thread count: the configured `threads`, else the one host
  // thread that drives the device
  b["threads"] = Json((int64_t)std::max(1, s.thread_count()));
  b["solver_time_ms"] = t2 - t1;
  b["cpu_time_ms"] = cpu_ms() - c0;
  b["max_mem_bytes"] = Json(max_rss_bytes());
  b["disk_io_read_bytes"] = Json(rq.read_bytes);
  Json ext = b["extensions"];
  Json eb = Json::object();
  eb["name"] = "build";
  eb["value"] = build_properties();
  ext.push_back(eb);
  Json ei = Json::object();
  ei["name"] = "invocation";
  ei["value"] = invocation_properties();
  ext.push_back(ei);
  b["extensions"] = ext;
  b["end2end_time_ms"] = now_ms() - rq.t0_ms;
  b["end2end_cputime_ms"] = cpu_ms() - rq.cpu0_ms;
  b["preprocessing_time_ms"] = t1 - t0;
  b["postprocessing_time_ms"] = now_ms() - t2;
  // the response's own size, this field included (app/runner.cc:612-622)
  const std::string key = ",\"disk_io_write_bytes\":";
  const int64_t n0 = (int64_t)(r.dump().size() + key.size());
  int64_t n = n0 + 1;
  while (n != n0 + (int64_t)std::to_string(n).size()) n = n0 + (int64_t)std::to_string(n).size();
  b["disk_io_write_bytes"] = Json(n);
  return r;
}

template <class F>
int guarded(char** response, F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    return report(e, response);
  } catch (const JsonError& e) {
    // malformed text is a parsing error, a value of the wrong kind a type error
    const bool syntax = std::string(e.what()).rfind("JSON parse error", 0) == 0;
    return report(Error(syntax ? kParsingError : kInvalidTypes, e.what()), response);
  } catch (const std::exception& e) {
    return report(Error(kRuntimeError, e.what()), response);
  }
}

// Many solves of one target fused into shared device launches: configure
// and init every solve, plan the fusible ones into one batch, run it, read
// every fused result back in one transfer, then run the rest one by one.
struct BatchRun {
  std::vector<std::unique_ptr<Solver>> solvers;
  std::vector<int> fused;
  double init_ms = 0, run_ms = 0, finalize_ms = 0, device_ms = 0;
};

void run_batch(qiotk_engine e, const char* target, int32_t n, const qiotk_model* models,
               const char* const* params_json, BatchRun& br) {
  qmc_batch b = nullptr;
  device_check(qmc_batch_create(e->ctx, &b), "qmc_batch_create");
  std::unique_ptr<qmc_batch_s, void (*)(qmc_batch)> hold(b, qmc_batch_destroy);
  br.fused.assign(n, 0);
  const double t0 = now_ms();
  for (int32_t k = 0; k < n; ++k) {
    auto s = create_solver(target);
    s->set_model(&models[k]->model, models[k]->dev.get(), e->ctx);
    s->configure(Json::parse(params_json[k]));
    s->init_with_retry();
    br.fused[k] = s->batch_plan(b) ? 1 : 0;
    br.solvers.push_back(std::move(s));
  }
  const double t1 = now_ms();
  // every member finishes within its planned sweeps: one persistent launch
  // runs them all, or graph-replayed launches of the smallest requested
  // sweeps_per_launch
  int64_t spl = 0;
  for (int32_t k = 0; k < n; ++k)
    if (br.fused[k] && br.solvers[k]->sweeps_per_launch() > 0)
      spl = spl ? std::min(spl, br.solvers[k]->sweeps_per_launch()) : br.solvers[k]->sweeps_per_launch();
  if (spl > 0)
    device_check(qmc_batch_run_chunked(b, (int32_t)std::min<int64_t>(spl, 1 << 30)), "qmc_batch_run_chunked");
  else
    device_check(qmc_batch_run(b, 1 << 30), "qmc_batch_run");
  double ms = 0;
  device_check(qmc_batch_wait(b, &ms), "qmc_batch_wait");
  const double t2 = now_ms();
  int64_t tot_r = 0, tot_s = 0;
  for (int32_t k = 0; k < n; ++k) {
    const int64_t r = br.solvers[k]->batch_replicas();
    tot_r += r;
    tot_s += r * models[k]->model.n_vars;
  }
  std::vector<double> low(tot_r);
  std::vector<uint64_t> acc(tot_r);
  std::vector<uint8_t> st(tot_s);
  if (tot_r) device_check(qmc_batch_read(b, low.data(), acc.data(), st.data()), "qmc_batch_read");
  int64_t ro = 0, so = 0;
  for (int32_t k = 0; k < n; ++k) {
    Solver& s = *br.solvers[k];
    const int64_t r = s.batch_replicas();
    if (br.fused[k] && r) {
      s.finalize_from(low.data() + ro, acc.data() + ro, st.data() + so);
      ro += r;
      so += r * models[k]->model.n_vars;
    } else if (!br.fused[k]) {
      s.run();
      s.finalize();
    }
  }
  const double t3 = now_ms();
  br.init_ms = t1 - t0;
  br.run_ms = t2 - t1;
  br.finalize_ms = t3 - t2;
  br.device_ms = ms;
}

}  // namespace

extern "C" {

const char* qiotk_last_error(void) { return g_err.c_str(); }

// Runner::metric_to_console (app/runner.cc:640-688): the common, solver and
// problem metrics of a response's benchmark block between _QIO_METRICS labels.
int qiotk_metrics(const char* response_json, char** out) {
  return guarded(out, [&] {
    const Json r = Json::parse(response_json);
    if (!r.is_object() || !r.has("benchmark")) throw Error(kMissingInput, "the response has no benchmark block");
    const Json& b = r.at("benchmark");
    auto copy = [&](const char* k, Json& dst) { dst[k] = b.has(k) ? b.at(k) : Json((int64_t)0); };
    Json solver_props = Json::object();
    std::string model;
    if (b.has("extensions"))
      for (size_t i = 0; i < b.at("extensions").size(); ++i) {
        const Json& e = b.at("extensions")[i];
        if (e.at("name").as_string() == "solver") solver_props = e.at("value");
        if (e.at("name").as_string() == "device" && e.at("value").has("model"))
          model = e.at("value").at("model").as_string();
      }
    Json metrics = Json::object();
    Json common = Json::object();
    for (const char* k : {"solver_time_ms", "cpu_time_ms", "max_mem_bytes", "preprocessing_time_ms",
                          "postprocessing_time_ms"})
      copy(k, common);
    metrics["common"] = common;
    metrics["solver"] = b.has("solver") ? b.at("solver") : solver_props;
    Json problem = Json::object();
    for (const char* k : {"num_terms", "num_variables", "locality", "min_locality", "avg_locality",
                          "accumulated_dependent_vars", "sum_coefficient_degrees_total", "max_coupling_magnitude",
                          "min_coupling_magnitude"})
      copy(k, problem);
    problem["model"] = model;
    metrics["problem"] = problem;
    *out = dup("_QIO_METRICS" + metrics.dump() + "_QIO_METRICS");
    return 0;
  });
}

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
    m->model = read_model_text(input_json);
    m->model.source_bytes = (int64_t)std::strlen(input_json);
    if (!m->model.empty()) {
      m->dev = std::make_unique<DeviceModel>(e->ctx, m->model);
      m->dev->handle(QMC_VISIT_COLOR_CLASSES);  // resident: solves exclude the upload
    }
    *out = m.release();
    return 0;
  });
}

int qiotk_model_create_file(qiotk_engine e, const char* path, qiotk_model* out) {
  return guarded(nullptr, [&] {
    if (!path) throw Error(kMissingInput, "no input file");
    auto m = std::make_unique<qiotk_model_s>();
    m->e = e;
    m->model = read_model_file(path);
    m->model.source_bytes = path_bytes(path);
    if (!m->model.empty()) {
      m->dev = std::make_unique<DeviceModel>(e->ctx, m->model);
      m->dev->handle(QMC_VISIT_COLOR_CLASSES);  // resident: solves exclude the upload
    }
    *out = m.release();
    return 0;
  });
}

int qiotk_model_create_terms(qiotk_engine e, int kind, int64_t n_terms, const int64_t* offsets,
                             const int32_t* ids, const double* coeffs, qiotk_model* out) {
  return guarded(nullptr, [&] {
    auto m = std::make_unique<qiotk_model_s>();
    m->e = e;
    m->model = make_model_flat(kind == 0 ? QMC_ISING : QMC_PUBO, n_terms, offsets, ids, coeffs);
    if (!m->model.empty()) {
      m->dev = std::make_unique<DeviceModel>(e->ctx, m->model);
      m->dev->handle(QMC_VISIT_COLOR_CLASSES);  // resident: solves exclude the upload
    }
    *out = m.release();
    return 0;
  });
}

void qiotk_model_destroy(qiotk_model m) { delete m; }

int qiotk_solve_model(qiotk_model m, const char* target, const char* params_json, char** response) {
  return guarded(response, [&] {
    auto s = create_solver(target);
    s->set_model(&m->model, m->dev.get(), m->e->ctx);
    Request rq;
    rq.read_bytes = (int64_t)std::strlen(params_json) + m->model.source_bytes;
    s->configure(Json::parse(params_json));
    Json r = run_one(*s, rq);
    *response = dup(r.dump());
    return 0;
  });
}

int qiotk_solve(qiotk_engine e, const char* target, const char* params_json, const char* input_json,
                char** response) {
  return guarded(response, [&] {
    Request rq;  // configure time and its CPU time count towards end2end
    qiotk_model m = nullptr;
    int rc = qiotk_model_create(e, input_json, &m);
    if (rc) throw Error(rc, g_err);
    std::unique_ptr<qiotk_model_s> hold(m);
    rq.read_bytes = (int64_t)std::strlen(params_json) + m->model.source_bytes;
    auto s = create_solver(target);
    s->set_model(&m->model, m->dev.get(), e->ctx);
    s->configure(Json::parse(params_json));
    Json r = run_one(*s, rq);
    *response = dup(r.dump());
    return 0;
  });
}

int qiotk_solve_file(qiotk_engine e, const char* target, const char* params_json, const char* input_path,
                     char** response) {
  return guarded(response, [&] {
    Request rq;  // configure time and its CPU time count towards end2end
    qiotk_model m = nullptr;
    int rc = qiotk_model_create_file(e, input_path, &m);
    if (rc) throw Error(rc, g_err);
    std::unique_ptr<qiotk_model_s> hold(m);
    rq.read_bytes = (int64_t)std::strlen(params_json) + m->model.source_bytes;
    auto s = create_solver(target);
    s->set_model(&m->model, m->dev.get(), e->ctx);
    s->configure(Json::parse(params_json));
    Json r = run_one(*s, rq);
    *response = dup(r.dump());
    return 0;
  });
}

int qiotk_solve_batch(qiotk_engine e, const char* target, int32_t n, const qiotk_model* models,
                      const char* const* params_json, char** responses, double* device_ms) {
  for (int32_t k = 0; k < n; ++k) responses[k] = nullptr;
  return guarded(nullptr, [&] {
    BatchRun br;
    run_batch(e, target, n, models, params_json, br);
    for (int32_t k = 0; k < n; ++k) {
      Json r = br.solvers[k]->get_result();
      r["benchmark"]["fused"] = Json((bool)br.fused[k]);
      r["benchmark"]["batch_init_ms"] = br.init_ms;
      r["benchmark"]["batch_run_ms"] = br.run_ms;
      r["benchmark"]["batch_finalize_ms"] = br.finalize_ms;
      r["benchmark"]["batch_device_ms"] = br.device_ms;
      responses[k] = dup(r.dump());
    }
    if (device_ms) *device_ms = br.device_ms;
    return 0;
  });
}

int qiotk_solve_batch_arrays(qiotk_engine e, const char* target, int32_t n, const qiotk_model* models,
                             const char* const* params_json, double* cost, int8_t* configs,
                             double* timings) {
  return guarded(nullptr, [&] {
    BatchRun br;
    run_batch(e, target, n, models, params_json, br);
    int64_t off = 0;
    for (int32_t k = 0; k < n; ++k) {
      const int32_t nv = models[k]->model.n_vars;
      br.solvers[k]->best(cost ? cost + k : nullptr, configs ? configs + off : nullptr);
      off += nv;
    }
    if (timings) {
      timings[0] = br.init_ms;
      timings[1] = br.run_ms;
      timings[2] = br.finalize_ms;
      timings[3] = br.device_ms;
    }
    return 0;
  });
}

int qiotk_tune_batch(qiotk_engine e, const char* target, int32_t n, const qiotk_model* models,
                     const char* const* params_json, int32_t trial_sweeps, char** response) {
  return guarded(response, [&] {
    if (n <= 0) throw Error(kValueError, "nothing to tune");
    if (trial_sweeps < 2) throw Error(kValueError, "trial_sweeps must be >= 2");
    // sample: every k-th solve, at most 8192 -- the whole batch as a rule:
    // how its graphs share the L2 and the memory behind it decides the
    // shape, and a small sample fits caches the full batch does not
    const int32_t stride = std::max<int32_t>(1, (n + 8191) / 8192);
    std::vector<int32_t> idx;
    for (int32_t k = 0; k < n; k += stride) idx.push_back(k);
    std::vector<qiotk_model> ms;
    std::vector<Json> base;
    int64_t min_group = 32;  // replicas of the smallest run (SA restarts, PT ladder)
    const bool pt = std::string(target) == "paralleltempering.qiotoolkit";
    for (int32_t k : idx) {
      ms.push_back(models[k]);
      Json doc = Json::parse(params_json[k]);
      if (!doc.has("params")) throw Error(kMissingInput, "Input field `params` is required.");
      base.push_back(doc);
      auto s = create_solver(target);
      s->set_model(&models[k]->model, models[k]->dev.get(), e->ctx);
      s->configure(doc);
      min_group = std::min<int64_t>(min_group, s->run_replicas());
    }
    // one trial of the sample: replicas per work-group B (B_large for solves
    // of at least large_min variables), lane depth, sweeps per launch
    Json trials = Json::array();
    auto trial = [&](int B, int depth, int64_t large_min, int B_large, int depth_large, int64_t spl) {
      std::vector<std::string> ps;
      for (size_t i = 0; i < base.size(); ++i) {
        Json doc = base[i];
        Json& p = doc["params"];
        p["sweeps"] = Json((int64_t)trial_sweeps);
        p.remove("step_limit");
        const bool large = large_min > 0 && (int64_t)ms[i]->model.n_vars >= large_min;
        p["replicas_per_group"] = Json((int64_t)(large ? B_large : B));
        p["positions_per_lane"] = Json((int64_t)(large ? depth_large : depth));
        if (spl) p["sweeps_per_launch"] = Json(spl);
        ps.push_back(doc.dump());
      }
      std::vector<const char*> pp;
      for (auto& x : ps) pp.push_back(x.c_str());
      BatchRun br;
      run_batch(e, target, (int32_t)ms.size(), ms.data(), pp.data(), br);
      int64_t att = 0;
      for (auto& s : br.solvers) att += s->attempts();
      const double rate = br.device_ms > 0 ? (double)att / (br.device_ms * 1e-3) : 0.0;
      Json t = Json::object();
      t["replicas_per_group"] = Json((int64_t)B);
      t["positions_per_lane"] = Json((int64_t)depth);
      if (large_min > 0) {
        t["large_min_vars"] = Json(large_min);
        t["large_replicas_per_group"] = Json((int64_t)B_large);
        t["large_positions_per_lane"] = Json((int64_t)depth_large);
      }
      if (spl) t["sweeps_per_launch"] = Json(spl);
      t["device_ms"] = br.device_ms;
      t["attempts_per_s"] = rate;
      trials.push_back(t);
      return rate;
    };
    // 1. one shape for every solve: replicas per work-group x lane depth (0 =
    //    library default)
    double best_rate = -1;
    int best_B = 0, best_depth = 0;
    struct Cand {
      double rate;
      int B, depth;
    };
    std::vector<Cand> cands;
    for (int B : {1, 2, 4, 8, 16, 32}) {
      if (B > min_group) break;
      if (pt && min_group % B) continue;
      for (int depth : {0, 2, 8}) cands.push_back({trial(B, depth, 0, 0, 0, 0), B, depth});
    }
    // one short trial each is noisy (a few per cent): the shapes within 10 %
    // of the fastest (at most 3) are timed twice more and ranked by their
    // median
    std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.rate > b.rate; });
    for (size_t i = 0; i < cands.size() && i < 3; ++i) {
      if (cands[i].rate < 0.9 * cands[0].rate) break;
      double r[3] = {cands[i].rate, trial(cands[i].B, cands[i].depth, 0, 0, 0, 0),
                     trial(cands[i].B, cands[i].depth, 0, 0, 0, 0)};
      std::sort(r, r + 3);
      if (r[1] > best_rate) {
        best_rate = r[1];
        best_B = cands[i].B;
        best_depth = cands[i].depth;
      }
    }
    // a second shape for the largest solves, or launches in chunks, only
    // when they beat one persistent shape by more than the trials' noise
    const double margin = 1.03;
    // 2. the largest solves in another shape: a batch's largest solves run
    //    longest, and with many replicas per work-group they finish as one
    //    work-group per solve on a mostly idle device -- fewer replicas per
    //    work-group (more work-groups) or a shallower lane depth (wider
    //    work-groups) for the solves of at least the sample's 75th, 90th or
    //    97th percentile size
    int64_t best_large = 0;
    int best_B_large = 0, best_depth_large = 0;
    {
      std::vector<int64_t> sizes;
      for (auto m : ms) sizes.push_back((int64_t)m->model.n_vars);
      std::sort(sizes.begin(), sizes.end());
      std::vector<int64_t> thresholds;
      for (double q : {0.75, 0.9, 0.97}) {
        const size_t at = std::min(sizes.size() - 1, (size_t)(q * (double)sizes.size()));
        const int64_t thr = sizes[at];
        if (thr > sizes.front() && (thresholds.empty() || thresholds.back() != thr)) thresholds.push_back(thr);
      }
      std::vector<std::pair<int, int>> shapes;  // (replicas per work-group, lane depth)
      for (int Bl = best_B / 2; Bl >= 1 && Bl >= best_B / 4; Bl /= 2)
        if (!pt || min_group % Bl == 0) shapes.emplace_back(Bl, best_depth);
      for (int dl : {0, 2})
        if (dl != best_depth) shapes.emplace_back(best_B, dl);
      for (int64_t thr : thresholds)
        for (auto [Bl, dl] : shapes) {
          const double rate = trial(best_B, best_depth, thr, Bl, dl, 0);
          if (rate > best_rate * (best_large ? 1.0 : margin)) {
            best_rate = rate;
            best_large = thr;
            best_B_large = Bl;
            best_depth_large = dl;
          }
        }
    }
    // 3. sweeps per launch: one persistent launch (0) against graph-replayed
    //    chunks, for the chosen shape (split PT ladders always run persistent)
    int64_t best_spl = 0;
    if (!pt || best_B == min_group) {
      for (int64_t spl : {(int64_t)std::max(1, trial_sweeps / 4), (int64_t)std::max(1, trial_sweeps / 2)}) {
        const double rate = trial(best_B, best_depth, best_large, best_B_large, best_depth_large, spl);
        if (rate > best_rate * (best_spl ? 1.0 : margin)) {
          best_rate = rate;
          best_spl = spl;
        }
      }
    }
    Json r = Json::object();
    Json chosen = Json::object();
    chosen["replicas_per_group"] = Json((int64_t)best_B);
    chosen["positions_per_lane"] = Json((int64_t)best_depth);
    // a chunk length chosen on a short trial scales with the full run
    chosen["sweeps_per_launch"] = Json(best_spl);
    chosen["sweeps_per_launch_fraction"] = best_spl ? (double)best_spl / trial_sweeps : 0.0;
    // solves of at least large_min_vars variables take
    // large_replicas_per_group instead (0: one shape for all)
    chosen["large_min_vars"] = Json(best_large);
    chosen["large_replicas_per_group"] = Json((int64_t)best_B_large);
    chosen["large_positions_per_lane"] = Json((int64_t)best_depth_large);
    r["params"] = chosen;
    r["sample"] = Json((int64_t)idx.size());
    r["trial_sweeps"] = Json((int64_t)trial_sweeps);
    r["trials"] = trials;
    *response = dup(r.dump());
    return 0;
  });
}

int qiotk_model_vars(qiotk_model m, int32_t* n_vars, int32_t* names) {
  return guarded(nullptr, [&] {
    if (!m) throw Error(kValueError, "null model");
    if (n_vars) *n_vars = m->model.n_vars;
    if (names) std::memcpy(names, m->model.names.data(), m->model.names.size() * sizeof(int32_t));
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

static Json describe_model(const Model& m) {
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
  j["input_reader"] = m.streamed ? "stream" : "tree";
  if (!m.empty()) {
    j["estimate_max_cost_diff"] = m.estimate_max_cost_diff();
    j["estimate_min_cost_diff"] = m.estimate_min_cost_diff();
  }
  Json ic = Json::array();
  for (auto& kv : m.initial) {
    Json e = Json::array();
    e.push_back(Json((int64_t)kv.first));
    e.push_back(Json((int64_t)kv.second));
    ic.push_back(e);
  }
  j["initial"] = ic;
  return j;
}

int qiotk_parse_model(const char* input_json, char** response) {
  return guarded(response, [&] {
    *response = dup(describe_model(read_model_text(input_json)).dump());
    return 0;
  });
}

int qiotk_parse_model_file(const char* path, char** response) {
  return guarded(response, [&] {
    if (!path) throw Error(kMissingInput, "no input file");
    *response = dup(describe_model(read_model_file(path)).dump());
    return 0;
  });
}

int qiotk_observe(const char* params_json, const char* events_json, char** response) {
  return guarded(response, [&] {
    Observer o;
    o.configure(Json::parse(params_json));
    const Json ev = Json::parse(events_json);
    if (!ev.is_array()) throw Error(kInvalidTypes, "events must be an array");
    for (const Json& e : ev.items()) {
      if (e.has("label")) {
        const Json& v = e.at("value");
        if (v.is_null()) o.clear_label(e.at("label").as_string());
        else o.set_label(e.at("label").as_string(), v.as_int());
      } else if (e.has("restart")) {
        o.restart(e.at("restart").as_int());
      } else {
        o.observe(e.at("observe").as_string(), e.at("value").as_double(),
                  e.has("weight") ? e.at("weight").as_double() : 1.0);
      }
    }
    *response = dup(o.render().dump());
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
