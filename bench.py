#!/usr/bin/env python3
"""Benchmark of the MI355X sweep engine.

The default run (``python bench.py``) prints one JSON line.  Its top level is
parallel tempering on a 16^3 periodic cubic spin glass (N = 4096 spins, 12288
couplings J ~ N(0,1) fp32, h = 0), 32 replicas at geometric temperatures
0.2 -> 2.0, replica exchange after every sweep, 5000 sweeps, 16 independent
restarts (seeds s..s+15).  One step = the whole job: 16 solves configured,
fused into one device launch and read back (best cost and best
configuration of each, as arrays), with the problem already resident in
HBM.  Metric: spin-flip attempts/s (= restarts x replicas x sweeps x N / s of
wall time).  Nested under their own keys, in the same run:

  c4  the uneven batch: 4096 ER instances, 8 restarts x 1000 SA sweeps each,
      autotuned shared launches against the plain one-instance-per-launch form
  c5  the serving stream: Poisson arrivals of mixed Ising / PUBO requests,
      population annealing or SA, p50 / p99 latency and the sustained rate,
      offered at 6 requests/s (the highest rate the loop keeps up with on
      one MI355X: over 60 s at 6 requests/s at most 2 requests wait for
      admission and the latency stays flat; at 7 the device is 86 % busy)
  c1  SA on the 32x32 +-1 lattice, 64 restarts x 1000 sweeps
  c3  SA on the 2000-variable random PUBO, 256 restarts x 2000 sweeps

Each carries a roofline object for its dominant kernel: the algorithmic bytes
per attempt B = d*(4 + s_c + 1) + 2 (d incident terms, s_c coupling bytes)
over the kernel's average launch time (HIP events on the stream the kernel
runs on), against the 8 TB/s HBM3E peak; ``bound`` names what the counters
show limits the kernel, ``traffic`` the HBM bytes per launch the counters
measured (profiles/).  The CPU baseline times the reference algorithm
restated in C++ (pcg32 streams, ln-table accept, the same PT loop), built on
this host with the reference's Release flags, on a bounded sample of the
same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# LDS: 256 CUs x 256 B/clk (CDNA4 LDS reads) x 2.4 GHz.  The 128 B/clk of
# earlier CDNA parts is exceeded on the device: a streaming 16-byte LDS read
# kernel (--lds-probe) measures 106-109 TB/s.
LDS_PEAK_GBS = 256 * 256 * 2.4

# HBM bytes per step of the default shapes (c4: the three sweep launches of a
# step; the others: the one sweep launch), from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE in separate runs, in KiB, scaled by what the same counters
# report for a plain 512 MiB copy (FETCH_SIZE x 2, WRITE_SIZE x 1).
# Summaries: profiles/k_sweep_*.md
MEASURED_TRAFFIC_BYTES = {
    "c2": int((107108.3 * 2 + 82106.0) * 1024),
    "c1": int((328.4 * 2 + 142.0) * 1024),
    "c3": int((6100.3 * 2 + 1056.3) * 1024),
    "c4": int((107989455.8 * 2 + 9075.9) * 1024),
}

# What limits each kernel (rocprofv3 kernel traces, SQ and byte counters in
# profiles/); frac prices the algorithmic bytes against the HBM peak, above 1
# where on-chip reuse serves more bytes than HBM could.
_BASIS = "; frac = algorithmic bytes over the HBM peak (not an HBM bound)"
LIMITER = {
    "c2": "on-chip: VALU issue (0.59 of the issue floor) and the cross-work-group ladder exchange; "
          "graph in L2, state in LDS" + _BASIS,
    "c1": "per-class latency (wave scans, barrier, combine: 2 classes of one round) with 64 work-groups "
          "for 256 CUs; graph and state on chip" + _BASIS,
    "c3": "per-class latency (20 colour classes of one 512-lane round each) and VALU issue of the "
          "4-local fields; graph in L2, state in LDS" + _BASIS,
    "c4": "the largest instances' work-groups (the last of three concurrent launches) and their graph "
          "stream from memory (the batch's graphs exceed L2)" + _BASIS,
}


def limiter(name):
    return LIMITER[name]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def roofline(attempts, B, kernel_ms, kernel, traffic=None, workload="c2"):
    achieved = attempts * B / (kernel_ms * 1e-3) / 1e9
    return {"bound": limiter(workload), "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_attempt": B,
            "kernel": kernel, "kernel_ms": kernel_ms, "lds_frac": achieved / LDS_PEAK_GBS}


def c2_workload(seed):
    from qiotoolkit_amd import workloads as wl
    terms = wl.lattice3d(16, seed=seed)
    inp = {"cost_function": {"type": "ising", "version": "1.0", "terms": terms}}
    temps = {"type": "geometric", "initial": 0.2, "final": 2.0, "count": 32}
    return inp, temps


def run_c2(args):
    from qiotoolkit_amd.solver import Engine
    restarts, sweeps, L, N = 16, args.sweeps, 32, 4096
    inp, temps = c2_workload(20241203)
    eng = Engine(0)
    model = eng.model(inp)
    params = []
    for k in range(restarts):
        p = {"temperatures": temps, "sweeps": sweeps, "sweeps_per_replica": 1, "seed": 1000 + k}
        if args.group:
            p["replicas_per_group"] = args.group
        if args.lanes:
            p["lanes"] = args.lanes
        params.append({"params": p})
    attempts_per_step = restarts * L * sweeps * N

    def step():
        # results as arrays: best cost and best configuration of every restart
        return eng.solve_batch_arrays("paralleltempering.qiotoolkit", [model] * restarts, params)

    try:
        for _ in range(args.warmup):
            step()
        t0 = time.perf_counter()
        dev_ms, best = [], None
        for _ in range(args.steps):
            costs, configs, tm = step()
            dev_ms.append(tm["device_ms"])
            best = list(costs)
        t = time.perf_counter() - t0
    finally:
        model.close()
        eng.close()
    log("phases ms: init %.1f run %.1f finalize %.1f device %.1f" % (
        tm["init_ms"], tm["run_ms"], tm["finalize_ms"], tm["device_ms"]))
    value = attempts_per_step * args.steps / t
    avg_ms = float(np.mean(dev_ms))
    B = 6 * (4 + 4 + 1) + 2
    default_shape = args.sweeps == 5000 and not args.group and not args.lanes
    traffic = MEASURED_TRAFFIC_BYTES["c2"] if default_shape else None
    out = {
        "metric": "spin-flip attempts/s at equal final-energy quality; p50/p99 solve latency",
        "value": value,
        "unit": "spin-flip attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 16^3 spin glass, J~N(0,1) rounded to fp32)",
        "config": {"workload": "PT 3D spin glass 16^3 periodic, 32 replicas geometric T 0.2->2.0, "
                               "exchange every sweep, %d sweeps, 16 restarts" % sweeps,
                   "n_spins": N, "couplings": 12288, "replicas": L, "restarts": restarts,
                   "sweeps": sweeps, "parallelism": "none (1 GPU)"},
        "roofline": roofline(attempts_per_step, B, avg_ms, "k_sweep<ISING,pairwise,f32>", traffic),
        "quality": {"best_energy_mean": float(np.mean(best)), "best_energy_min": float(np.min(best)),
                    "per_spin_mean": float(np.mean(best)) / N},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_c2(inp, temps, args)
    return out


SA_WORKLOADS = {
    # name: (kind, terms, params, incident terms per variable, coupling bytes, dtype, description)
    "c1": ("ising", lambda wl: wl.lattice2d(32, seed=1),
           {"beta_start": 0.1, "beta_stop": 3.0, "restarts": 64, "sweeps": 1000}, 4, 4, "int32",
           "SA 2D square lattice 32x32 periodic, J in {-1,+1}, h = 0, 64 restarts x 1000 sweeps, beta 0.1->3.0"),
    "c3": ("pubo", lambda wl: wl.random_pubo(2000, 20000, seed=1),
           {"beta_start": 0.01, "beta_stop": 5.0, "restarts": 256, "sweeps": 2000}, 26, 4, "int32",
           "SA random PUBO 2000 vars, 20000 terms of degree 1/2/3/4 at 10/40/30/20%, c in [-10,10]\\{0}, "
           "256 restarts x 2000 sweeps, beta 0.01->5.0"),
}


def run_sa(args, name, steps, warmup):
    """One simulated-annealing solve per step on a problem resident in HBM."""
    from qiotoolkit_amd import workloads as wl
    from qiotoolkit_amd.solver import Engine
    kind, terms, params, d, sc, dtype, desc = SA_WORKLOADS[name]
    params = dict(params)
    if args.group:
        params["replicas_per_group"] = args.group
    if args.lanes:
        params["lanes"] = args.lanes
    eng = Engine(0)
    model = eng.model({"cost_function": {"type": kind, "version": "1.0", "terms": terms(wl)}})
    n = model.n_vars()
    target = "simulatedannealing.qiotoolkit"
    plist = [{"params": dict(params, seed=300)}]
    try:
        for _ in range(warmup):
            eng.solve_batch_arrays(target, [model], plist, configs=False)
        dev_ms, best = [], []
        t0 = time.perf_counter()
        for s in range(steps):
            cost, _, tm = eng.solve_batch_arrays(target, [model], [{"params": dict(params, seed=300 + s)}],
                                                 configs=True)
            dev_ms.append(tm["device_ms"])
            best.append(float(cost[0]))
        t = time.perf_counter() - t0
    finally:
        model.close()
        eng.close()
    attempts = params["restarts"] * params["sweeps"] * n
    B = d * (4 + sc + 1) + 2
    default_shape = not args.group and not args.lanes
    return {"value": attempts * steps / t, "unit": "spin-flip attempts/s", "steps": steps, "warmup": warmup,
            "ms_per_step": t * 1e3 / steps, "dtype": dtype,
            "config": {"workload": desc, "n_vars": n, "attempts_per_step": attempts},
            "roofline": roofline(attempts, B, float(np.mean(dev_ms)), "k_sweep (device time of the solve)",
                                 MEASURED_TRAFFIC_BYTES[name] if default_shape else None, name),
            "quality": {"best_cost_mean": float(np.mean(best)), "best_cost_min": float(np.min(best))}}


def c4_instances(n=4096, seed=4096):
    from qiotoolkit_amd import workloads as wl
    return wl.batched_er_instances(n, seed)


def run_c4(args):
    """Batched uneven SA: every instance 8 restarts x `sweeps` sweeps, linear
    beta 0.1 -> 3.0, fused into shared launches, against the same solves run
    one instance per launch (a sample of them) -- same seeds, same results."""
    from qiotoolkit_amd.solver import Engine
    sweeps = args.sweeps if args.sweeps != 5000 else 1000
    t0 = time.perf_counter()
    insts = c4_instances(args.instances)
    eng = Engine(0)
    models = [eng.model_from_arrays("ising", off, ids, c) for (_, off, ids, c) in insts]
    log("c4: %d instances, %d spins, ingest %.1f s" % (len(insts), sum(i[0] for i in insts),
                                                       time.perf_counter() - t0))
    extra = {}
    if args.group:
        extra["replicas_per_group"] = args.group
    if args.lanes:
        extra["lanes"] = args.lanes
    if args.positions_per_lane:
        extra["positions_per_lane"] = args.positions_per_lane
    target = "simulatedannealing.qiotoolkit"

    def params(k):
        p = {"beta_start": 0.1, "beta_stop": 3.0, "restarts": 8, "sweeps": sweeps, "seed": 7000 + k}
        p.update(extra)
        return {"params": p}

    plist = [params(k) for k in range(len(models))]
    tune = None
    if not (args.group or args.lanes or args.positions_per_lane or args.no_tune):
        t1 = time.perf_counter()
        tune = eng.tune_batch(target, models, plist, trial_sweeps=20)
        tune["seconds"] = time.perf_counter() - t1
        ch = tune["params"]
        extra["replicas_per_group"] = ch["replicas_per_group"]
        extra["positions_per_lane"] = ch["positions_per_lane"]
        if ch["sweeps_per_launch_fraction"] > 0:
            extra["sweeps_per_launch"] = max(1, int(round(ch["sweeps_per_launch_fraction"] * sweeps)))
        log("c4 autotune %.1f s: %s" % (tune["seconds"], json.dumps(ch)))
        plist = [params(k) for k in range(len(models))]
        if ch.get("large_min_vars", 0) > 0:
            # the largest solves spread over more work-groups
            for k, inst in enumerate(insts):
                if inst[0] >= ch["large_min_vars"]:
                    plist[k]["params"]["replicas_per_group"] = ch["large_replicas_per_group"]
                    plist[k]["params"]["positions_per_lane"] = ch["large_positions_per_lane"]
            for key in ("large_min_vars", "large_replicas_per_group", "large_positions_per_lane"):
                extra[key] = ch[key]
    attempts = 8 * sweeps * sum(i[0] for i in insts)
    # the parameter documents are inputs: encoded once, like the models
    penc = [json.dumps(p).encode() for p in plist]
    for _ in range(args.warmup):
        eng.solve_batch_arrays(target, models, penc, configs=False)
    t0 = time.perf_counter()
    dev_ms, tms = [], None
    for _ in range(args.steps):
        costs, cfg, tms = eng.solve_batch_arrays(target, models, penc, configs=True)
        dev_ms.append(tms["device_ms"])
    t = time.perf_counter() - t0
    log("c4 phases ms: init %.1f run %.1f finalize %.1f device %.1f" % (
        tms["init_ms"], tms["run_ms"], tms["finalize_ms"], tms["device_ms"]))
    value = attempts * args.steps / t
    avg_ms = float(np.mean(dev_ms))
    B = 7 * (4 + 1 + 1) + 2   # d = 6 couplings + 1 field term, int8 coefficients
    # the profiled shape: the tuner's usual choice, 8 replicas per work-group
    # and lane depth 8, one tier, one launch
    default_shape = (sweeps == 1000 and len(models) == 4096 and tune is not None
                     and extra.get("replicas_per_group") == 8 and extra.get("positions_per_lane") == 8
                     and not extra.get("large_min_vars") and "sweeps_per_launch" not in extra)
    # plain form: the same solves, one instance per launch, on a sample --
    # with the batch's parameters, and with one replica per work-group (the
    # plain form's best shape; the grouping does not change results)
    stride = max(1, len(models) // max(1, args.plain_sample))
    idx = list(range(0, len(models), stride)) if args.plain_sample > 0 else []
    plain = {}
    same = True
    for label, over in (("batch_params", {}), ("one_replica_per_group", {"replicas_per_group": 1})):
        plain_t, plain_att = 0.0, 0
        for k in idx:
            pk = {"params": dict(plist[k]["params"], **over)}
            pk["params"].pop("sweeps_per_launch", None)
            t1 = time.perf_counter()
            c1, _, _ = eng.solve_batch_arrays(target, [models[k]], [pk], configs=False)
            plain_t += time.perf_counter() - t1
            plain_att += 8 * sweeps * insts[k][0]
            same &= bool(c1[0] == costs[k])
        if plain_t > 0:
            plain[label] = plain_att / plain_t
    plain_best = max(plain.values()) if plain else None
    for m in models:
        m.close()
    eng.close()
    out = {
        "metric": "spin-flip attempts/s at equal final-energy quality; p50/p99 solve latency",
        "value": value,
        "unit": "spin-flip attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded Erdos-Renyi instances)",
        "config": {"workload": "batched uneven SA: %d ER instances N log-uniform [64,4096], mean degree 6, "
                               "J int in [-3,3]\\{0}, h in {-1,0,1}, 8 restarts x %d sweeps, beta 0.1->3.0"
                               % (len(models), sweeps),
                   "instances": len(models), "spins": sum(i[0] for i in insts), "restarts": 8,
                   "sweeps": sweeps, "params": extra, "parallelism": "none (1 GPU)"},
        "roofline": roofline(attempts, B, avg_ms, "k_sweep<ISING,pairwise,i32> (all launches of a step)",
                             MEASURED_TRAFFIC_BYTES["c4"] if default_shape else None, "c4"),
        "plain_per_instance": {"value": plain_best, "by_shape": plain, "instances_sampled": len(idx),
                               "speedup_of_batched": value / plain_best if plain_best else None,
                               "identical_best_costs": same},
        "autotune": tune,
        "quality": {"best_cost_mean": float(np.mean(costs))},
    }
    return out


def c5_requests(rate, duration, seed=5005, pa_extra=None):
    from qiotoolkit_amd import workloads as wl
    return wl.request_stream(rate, duration, seed, pa_extra)


def run_c5(args):
    """Serving stream: requests submitted at their Poisson arrival times to
    the serving loop; latency = finish - scheduled arrival."""
    from qiotoolkit_amd.solver import Server
    t_gen = time.perf_counter()
    pa_extra = {}
    if args.pa_group:
        pa_extra["replicas_per_group"] = args.pa_group
    if args.pa_order:
        pa_extra["sweep_order"] = args.pa_order
    if args.pa_lanes:
        pa_extra["lanes"] = args.pa_lanes
    reqs = c5_requests(args.rate, args.duration, pa_extra=pa_extra)
    log("c5: %d requests generated in %.1f s (%d ising, %d pubo, %d PA, %d SA)" % (
        len(reqs), time.perf_counter() - t_gen, sum(r["kind"] == "ising" for r in reqs),
        sum(r["kind"] == "pubo" for r in reqs), sum(r["target"].startswith("pop") for r in reqs),
        sum(r["target"].startswith("sim") for r in reqs)))
    cfg = {"max_active": args.max_active, "sweeps_per_round": args.sweeps_per_round}
    if args.hbm_budget_gib:
        cfg["hbm_budget_bytes"] = int(args.hbm_budget_gib * (1 << 30))
    srv = Server(0, **cfg)
    t0 = time.perf_counter()
    ids, lag = [], 0.0
    for r in reqs:
        now = time.perf_counter() - t0
        if r["t"] > now:
            time.sleep(r["t"] - now)
        lag = max(lag, time.perf_counter() - t0 - r["t"])
        ids.append(srv.submit(r["target"], {"params": r["params"]}, r["kind"], *r["arrays"]))
    log("c5: all submitted at %.1f s (max submit lag %.3f s)" % (time.perf_counter() - t0, lag))
    lat, att, fails = [], 0, 0
    for r, rid in zip(reqs, ids):
        rc, doc = srv.wait(rid, -1)
        if rc:
            fails += 1
            continue
        sv = doc["benchmark"]["serving"]
        lat.append(sv["finished_ms"] - r["t"] * 1e3)
        att += doc["benchmark"]["cost_difference_evaluation_count"]
    t_end = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array(lat)
    value = att / t_end
    out = {
        "metric": "spin-flip attempts/s at equal final-energy quality; p50/p99 solve latency",
        "value": value,
        "unit": "spin-flip attempts/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": t_end * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/int32",
        "data": "synthetic (seeded request stream)",
        "config": {"workload": "serving stream: Poisson arrivals %.2f/s for %.0f s; 70%% Ising ER N 100-50000 "
                               "deg 8 fp32, 30%% PUBO 500-20000 vars 10x terms; PA (pop 512, 200 steps x 10 "
                               "sweeps, beta 0->3) or SA (64 x 1000 sweeps, beta 0.1->3) 50/50"
                               % (args.rate, args.duration),
                   "requests": len(reqs), "offered_rate_per_s": args.rate, "server": cfg,
                   "pa_params": pa_extra, "parallelism": "none (1 GPU)"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)) if len(lat) else None,
                       "p99": float(np.percentile(lat, 99)) if len(lat) else None,
                       "max": float(lat.max()) if len(lat) else None},
        "requests_per_s": len(lat) / t_end,
        # completed requests per second between the first arrival and the
        # last completion (the drain after the last arrival keeps it below
        # the offered rate even when the loop keeps up)
        "sustained_rate_per_s": len(lat) / max(1e-9, t_end - (reqs[0]["t"] if reqs else 0.0)),
        # the loop keeps up when the later arrivals wait no longer than the
        # earlier ones and few requests wait for admission
        "latency_p50_by_arrival_half_ms": [float(np.percentile(h, 50)) if len(h) else None
                                           for h in (lat[:len(lat) // 2], lat[len(lat) // 2:])],
        "most_waiting_for_admission": st.get("peak_pending"),
        "failed": fails,
        "server_stats": st,
    }
    return out


def _host_oracle():
    """Build the oracle for this host (-march=native, the reference's Release
    flags) and return its variant; the portable -march=x86-64-v3 build if the
    host build fails."""
    try:
        subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "liboracle_native.so"],
                       check=True, timeout=180, capture_output=True)
        return "_native", "-O3 -Ofast -march=native -mtune=native -funroll-loops -fopenmp"
    except Exception as e:  # noqa: BLE001
        log("host-tuned oracle build failed (%s); using the portable build" % e)
        return "_fast", "-O3 -Ofast -march=x86-64-v3 -mtune=generic -funroll-loops -fopenmp"


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_c2(inp, temps, args):
    """Reference algorithm (oracle) on a bounded sample: one 32-replica ladder."""
    from oracle import pyoracle as po
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(32, cores))
    variant, flags = _host_oracle()
    p = po.Problem("ising", inp["cost_function"]["terms"], variant=variant)
    tvals = [temps["initial"] * (temps["final"] / temps["initial"]) ** (i / 31) for i in range(32)]
    # calibrate on a short run, then whole ladders (restarts of the
    # workload, 5000 sweeps each) until the sample reaches about
    # args.cpu_seconds, at most the workload's 16
    o = po.make_opts(po.RNG_REFERENCE, seed=7, threads=threads, steps_per_tick=1)
    t0 = time.perf_counter()
    p.pt(o, tvals, False, 20)
    dt = max(1e-3, time.perf_counter() - t0)
    sweeps = 5000 if args.sweeps == 5000 else args.sweeps
    per_ladder = dt * sweeps / 20
    ladders = int(max(1, min(16, round(args.cpu_seconds / per_ladder))))
    if per_ladder > 2 * args.cpu_seconds:  # a slow host: one partial ladder
        sweeps, ladders = int(max(20, sweeps * args.cpu_seconds / per_ladder)), 1
    t0 = time.perf_counter()
    for k in range(ladders):
        p.pt(po.make_opts(po.RNG_REFERENCE, seed=7 + k, threads=threads, steps_per_tick=1), tvals, False, sweeps)
    dt = time.perf_counter() - t0
    attempts = ladders * 32 * 4096 * sweeps
    # one ladder on one thread, over a fifth of the sweeps (bounded)
    o1 = po.make_opts(po.RNG_REFERENCE, seed=7, threads=1, steps_per_tick=1)
    sweeps1 = max(20, sweeps // 5)
    t1 = time.perf_counter()
    p.pt(o1, tvals, False, sweeps1)
    dt1 = time.perf_counter() - t1
    single = 32 * 4096 * sweeps1 / dt1
    return {"value": attempts / dt, "unit": "spin-flip attempts/s", "cores": threads,
            "kind": "port",
            "single_thread": single,
            "thread_scaling": (attempts / dt) / single,
            "cpu_model": _cpu_model(),
            "build": "oracle/liboracle%s.so built on this host: %s" % (variant, flags),
            "sample": "%d of the 16 restarts of the same workload (32-replica ladders, %d sweeps each), %.1f s, "
                      "pcg32 + ln-table accept, OpenMP over replicas on %d threads; single_thread: one ladder, %d "
                      "sweeps on 1 thread, %.1f s" % (ladders, sweeps, dt, threads, sweeps1, dt1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="all",
                    help="all (c2 at the top level, c4 c5 c1 c3 nested), or one of c2 c4 c5 c1 c3")
    ap.add_argument("--sweeps", type=int, default=5000)
    ap.add_argument("--group", type=int, default=0, help="replicas per work-group (0 = auto)")
    ap.add_argument("--lanes", type=int, default=0, help="positions per class round (0 = auto)")
    ap.add_argument("--positions-per-lane", type=int, default=0,
                    help="lanes cover the largest colour class in this many rounds (0 = auto)")
    ap.add_argument("--instances", type=int, default=4096, help="c4: number of instances")
    ap.add_argument("--no-tune", action="store_true", help="c4: skip the autotuner")
    ap.add_argument("--rate", type=float, default=6.0, help="c5: offered requests per second")
    ap.add_argument("--duration", type=float, default=30.0, help="c5: seconds of arrivals")
    ap.add_argument("--max-active", type=int, default=64, help="c5: concurrently admitted requests")
    ap.add_argument("--sweeps-per-round", type=int, default=10, help="c5: sweeps per shared launch")
    ap.add_argument("--hbm-budget-gib", type=float, default=0.0, help="c5: device memory for requests")
    ap.add_argument("--pa-group", type=int, default=0, help="c5: replicas per work-group of PA requests (0 = solver default)")
    ap.add_argument("--pa-order", default="", help="c5: sweep_order of PA requests (default: the solver's, index)")
    ap.add_argument("--pa-lanes", type=int, default=0, help="c5: lanes of PA work-groups (0 = library default)")
    ap.add_argument("--plain-sample", type=int, default=128,
                    help="c4: instances run one per launch for the plain-form rate (0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates the HBM byte counters)")
    ap.add_argument("--lds-probe", action="store_true",
                    help="only measure the aggregate LDS read bandwidth (the LDS peak of the roofline)")
    args = ap.parse_args()
    if args.lds_probe:
        from qiotoolkit_amd import _qmc
        ctx = _qmc.Context(0)
        bw = [ctx.lds_bandwidth() for _ in range(3)]
        print(json.dumps({"lds_read_bytes_per_s": max(bw), "runs": bw, "assumed_peak": LDS_PEAK_GBS * 1e9}))
        ctx.close()
        return
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        for _ in range(3):
            y = x.clone()
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4,
                          "copies": 3, "check": float(y[-1].item())}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X; "
              "multi-GPU runs are not part of this build (--gpus %d refused)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    runs = {"c2": run_c2, "c4": run_c4, "c5": run_c5,
            "c1": lambda a: run_sa(a, "c1", a.steps, a.warmup), "c3": lambda a: run_sa(a, "c3", a.steps, a.warmup)}
    if args.workload == "all":
        out = run_c2(args)
        nested = [
            ("c4", lambda: run_c4(argparse.Namespace(**dict(vars(args), steps=min(args.steps, 3),
                                                             warmup=min(args.warmup, 1), sweeps=1000)))),
            ("c5", lambda: run_c5(args)),
            ("c1", lambda: run_sa(args, "c1", args.steps, args.warmup)),
            ("c3", lambda: run_sa(args, "c3", min(args.steps, 5), min(args.warmup, 1))),
        ]
        drop = ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline")
        for name, fn in nested:
            t0 = time.perf_counter()
            try:
                res = fn()
                out[name] = {k: v for k, v in res.items() if k not in drop}
            except Exception as e:  # noqa: BLE001 -- the line still reports the other workloads
                out[name] = {"error": "%s: %s" % (type(e).__name__, e)}
            log("%s done in %.1f s" % (name, time.perf_counter() - t0))
    elif args.workload in runs:
        out = runs[args.workload](args)
    else:
        print("unknown workload %s" % args.workload, file=sys.stderr)
        sys.exit(2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
