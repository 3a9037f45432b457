/* qiotk.h -- C ABI of the host solver library (libqiotk.so).
 *
 * The request/response surface of the reference's Runner and CLI
 * (app/runner.cc:35-429, 513-688; app/qiotoolkit.cc:128-222): a solver
 * target, its params JSON ({"params": {...}}) and the cost function JSON
 * ({"cost_function": {...}}), answered with {"benchmark": {...},
 * "solutions": {...}}.  Errors return the reference's codes
 * (utils/error_handling.h:25-52) with a message.
 *
 * Targets: simulatedannealing.qiotoolkit, paralleltempering.qiotoolkit,
 * populationannealing.cpu (app/runner.cc:137-148), and the parameter-free
 * simulatedannealing-parameterfree.qiotoolkit and
 * paralleltempering-parameterfree.qiotoolkit (app/runner.cc:172-181; not
 * taken by the serving loop).
 */
#ifndef QIOTK_H
#define QIOTK_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct qiotk_engine_s* qiotk_engine;
typedef struct qiotk_model_s* qiotk_model;

/* one engine per device; owns the device context */
int qiotk_engine_create(int device, qiotk_engine* out);
void qiotk_engine_destroy(qiotk_engine e);

/* One solve: Runner::configure + Runner::get_run_output.  *response is
 * malloc'ed JSON (free with qiotk_free); on error it holds
 * {"error": {"code": c, "message": m}}. */
int qiotk_solve(qiotk_engine e, const char* target, const char* params_json,
                const char* input_json, char** response);

/* A model parsed, normalised and resident on the device, reused by many
 * solves (timed solves then exclude ingest and upload). */
int qiotk_model_create(qiotk_engine e, const char* input_json, qiotk_model* out);
/* The same from a term list already in memory (kind 0 ising, 1 pubo; term k
 * has coefficient coeffs[k] and ids ids[offsets[k]..offsets[k+1])), as the
 * reference's protobuf ingest provides an alternative to JSON
 * (utils/proto_reader.h:196-228). */
int qiotk_model_create_terms(qiotk_engine e, int kind, int64_t n_terms, const int64_t* offsets,
                             const int32_t* ids, const double* coeffs, qiotk_model* out);
void qiotk_model_destroy(qiotk_model m);
int qiotk_solve_model(qiotk_model m, const char* target, const char* params_json, char** response);

/* Input files as Runner::configure reads them (app/runner.cc:215-260): a
 * JSON document, or a folder of protobuf problem messages
 * <folder>/<name>_0.pb, <name>_1.pb, ... (name = the folder's last path
 * component; utils/proto_reader.h:38-145, utils/problem.proto:1-30), read
 * until the next file is missing -- the first message gives the type and
 * version, every message adds terms.  Errors: the reference's messages with
 * FileReadError 301 (first file missing), MissingInput 102 (no cost
 * function), ValueError 107 (type, version, no terms). */
int qiotk_model_create_file(qiotk_engine e, const char* path, qiotk_model* out);
int qiotk_solve_file(qiotk_engine e, const char* target, const char* params_json, const char* input_path,
                     char** response);

/* Many independent solves of one target on resident models, fused into
 * shared device launches (simulated annealing and parallel tempering):
 * models[k] with params_json[k].  responses[k] as for qiotk_solve; *device_ms
 * = summed device time of the fused launches. */
int qiotk_solve_batch(qiotk_engine e, const char* target, int32_t n, const qiotk_model* models,
                      const char* const* params_json, char** responses, double* device_ms);
/* The same solves with results as arrays instead of JSON documents:
 * cost[k] = the response's solutions.cost; configs (may be NULL) receives
 * n_vars(k) values per solve, concatenated in k order, holding the
 * response's solutions.configuration by internal variable id (+1/-1 Ising,
 * 1/0 PUBO; names from qiotk_model_vars).  timings (may be NULL) = {host
 * init ms, run ms, finalize ms, device ms}. */
int qiotk_solve_batch_arrays(qiotk_engine e, const char* target, int32_t n, const qiotk_model* models,
                             const char* const* params_json, double* cost, int8_t* configs,
                             double* timings);
/* Autotuner for batched solves: short trials (trial_sweeps sweeps each, the
 * schedule compressed accordingly) of candidate launch shapes -- replicas per
 * work-group x lane depth (positions each lane covers per colour class,
 * which sets the LDS tile of a class round) -- on the solves (every k-th one
 * past 8192), timed with HIP events; then, for the fastest shape, another
 * shape for the largest solves (those of at least the sample's 75th, 90th
 * or 97th percentile size): 1/2 or 1/4 of the replicas per work-group (more
 * work-groups) or lane depth 0 or 2 (wider work-groups); then one
 * persistent launch against graph-replayed launches of 1/4 and 1/2 of the
 * trial sweeps.  Returned as
 * {"params": {"replicas_per_group": B, "positions_per_lane": k,
 * "large_min_vars": m (0 = none), "large_replicas_per_group": Bl,
 * "large_positions_per_lane": kl, "sweeps_per_launch": s (0 = persistent),
 * "sweeps_per_launch_fraction": s / trial_sweeps}, "trials": [...]}; merge
 * replicas_per_group and positions_per_lane (Bl and kl for solves of at
 * least m variables) into every solve's params and scale the fraction by
 * its sweeps.  replicas_per_group and sweeps_per_launch
 * never change results (the RNG counters see neither); positions_per_lane
 * changes the colour-class visit order. */
int qiotk_tune_batch(qiotk_engine e, const char* target, int32_t n, const qiotk_model* models,
                     const char* const* params_json, int32_t trial_sweeps, char** response);
/* Variable count and the original variable names in internal-id order
 * (first appearance in the term list, graph/graph.h:97-124); names may be NULL. */
int qiotk_model_vars(qiotk_model m, int32_t* n_vars, int32_t* names);

/* Serving loop (no counterpart in the reference, whose Runner solves one
 * request per process, app/runner.cc:513-688): a background thread with its
 * own device context admits submitted solves in arrival order while their
 * device footprint fits the HBM budget and a slot is free, then advances every
 * admitted solve in shared launches, round by round (sweeps_per_round sweeps
 * of each solve, or one population-annealing step), keeping its state on the
 * device until it completes.  Results are the qiotk_solve response plus
 * benchmark.serving {arrival_ms, admitted_ms, finished_ms, latency_ms,
 * queue_ms, rounds}; a result does not depend on what else was in flight.
 * config_json: {"hbm_budget_bytes": 68719476736, "max_active": 64,
 * "sweeps_per_round": 10, "round_target_ms": 2.0, "max_rounds_per_wait": 16}
 * (all optional): rounds are enqueued back to back until about
 * round_target_ms of device work is in flight (at most max_rounds_per_wait)
 * before the loop waits for the device, admits and completes solves. */
typedef struct qiotk_server_s* qiotk_server;
int qiotk_server_create(int device, const char* config_json, qiotk_server* out);
/* Enqueue a solve (model as a term CSR, kind 0 ising / 1 pubo, as in
 * qiotk_model_create_terms); returns at once with its request id. */
int qiotk_server_submit(qiotk_server s, const char* target, const char* params_json, int kind,
                        int64_t n_terms, const int64_t* offsets, const int32_t* ids, const double* coeffs,
                        int64_t* request_id);
int qiotk_server_submit_json(qiotk_server s, const char* target, const char* params_json,
                             const char* input_json, int64_t* request_id);
/* Wait up to timeout_ms (< 0: no limit) for a request; *response is NULL
 * when it is not finished yet, else its response (free with qiotk_free),
 * and the return value is its status code. */
int qiotk_server_wait(qiotk_server s, int64_t request_id, double timeout_ms, char** response);
/* Counters and latency percentiles of the finished requests as JSON. */
int qiotk_server_stats(qiotk_server s, char** response);
/* Finishes every submitted request, then stops the loop. */
void qiotk_server_destroy(qiotk_server s);
const char* qiotk_server_last_error(void);

/* Device-free host logic, for tooling and tests. */
/* graph::normalize_edges result as JSON: names, term CSR, const_cost, ...,
 * and for a non-empty model estimate_max_cost_diff / estimate_min_cost_diff
 * (model/ising.h:176-224, model/pubo.h:127-185), which set the
 * parameter-free beta range */
int qiotk_parse_model(const char* input_json, char** response);
/* the same for an input file (JSON document or protobuf problem folder) */
int qiotk_parse_model_file(const char* path, char** response);
/* Schedule::get_value at n points; *discretized = {"start","stop","count","values"} */
int qiotk_schedule_eval(const char* schedule_json, int32_t n, const double* x, double* out,
                        char** discretized);
/* The observer protocol (observe/observer.cc) of params_json's "observer"
 * fed with events_json, a list of {"label": name, "value": int | null},
 * {"observe": name, "value": v, "weight": w (default 1)} and {"restart":
 * step}; *response = its render, as benchmark.extensions "solver" holds it. */
int qiotk_observe(const char* params_json, const char* events_json, char** response);
/* Solver::configure only: the resolved plan (steps, betas, restarts) */
int qiotk_describe(const char* target, const char* params_json, char** response);

/* The _QIO_METRICS line the reference's CLI prints after a solve
 * (Runner::metric_to_console, app/runner.cc:640-688) for a response of
 * qiotk_solve*: "_QIO_METRICS{"common": {solver_time_ms, cpu_time_ms,
 * max_mem_bytes, preprocessing_time_ms, postprocessing_time_ms}, "solver":
 * {...}, "problem": {num_terms, num_variables, locality, min_locality,
 * avg_locality, accumulated_dependent_vars, sum_coefficient_degrees_total,
 * max_coupling_magnitude, min_coupling_magnitude, model}}_QIO_METRICS"
 * (absent numbers as 0).  Device-free. */
int qiotk_metrics(const char* response_json, char** out);

void qiotk_free(char* p);
const char* qiotk_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
