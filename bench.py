#!/usr/bin/env python3
"""Benchmark of the MI355X sweep engine on the parallel-tempering workload.

Default workload: parallel tempering on a 16^3 periodic cubic spin glass
(N = 4096 spins, 12288 couplings J ~ N(0,1) fp32, h = 0), 32 replicas at
geometric temperatures 0.2 -> 2.0, replica exchange after every sweep, 5000
sweeps, 16 independent restarts (seeds s..s+15).  One step = the whole job:
16 solves configured, fused into one device launch and read back (best cost
and best configuration of each, as arrays), with the problem already
resident in HBM.  Metric: spin-flip attempts/s (= restarts x replicas x
sweeps x N / s of wall time).

The roofline object prices the dominant kernel (the coloured sweep) with the
algorithmic bytes per attempt B = d*(4 + s_c + 1) + 2 (d = 6 incident terms,
s_c = 4 coupling bytes) = 56 B against the 8 TB/s HBM3E peak, measured live
from HIP events on the stream the kernel runs on.  The CPU baseline times the
reference algorithm restated in C++ (pcg32 streams, ln-table accept, the same
PT loop) on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
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

# HBM bytes per sweep-kernel launch of the default workload (5000 sweeps,
# automatic grouping and lanes), from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE in separate runs, scaled by what the same counters report for a
# plain 512 MiB copy (FETCH_SIZE x 2, WRITE_SIZE x 1):
# 205462.1 KiB x 2 + 82106 KiB.  Summary: profiles/k_sweep_pt_16cube_hbm_traffic.md
MEASURED_TRAFFIC_BYTES = {"c2": int((198183.25 * 2 + 82106) * 1024)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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
    achieved = attempts_per_step * B / (avg_ms * 1e-3) / 1e9
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_attempt": B, "kernel": "k_sweep<ISING,pairwise,f32>",
                     "kernel_ms": avg_ms, "lds_frac": achieved / LDS_PEAK_GBS},
        "quality": {"best_energy_mean": float(np.mean(best)), "best_energy_min": float(np.min(best)),
                    "per_spin_mean": float(np.mean(best)) / N},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_c2(inp, temps, args)
    return out


def c4_instances(n=4096, seed=4096):
    """4096 Erdos-Renyi instances, N log-uniform in [64, 4096], mean degree 6,
    J uniform in {-3..3}\\{0}, one field term h in {-1,0,1} per spin."""
    from qiotoolkit_amd import workloads as wl
    sizes = wl.log_uniform_sizes(n, 64, 4096, seed=seed)
    return [(int(N),) + wl.er_ising_arrays(int(N), 6.0, 3, seed=seed * 16 + k)
            for k, N in enumerate(sizes)]


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
        log("c4 autotune %.1f s: %s" % (tune["seconds"], json.dumps(extra)))
        plist = [params(k) for k in range(len(models))]
    attempts = 8 * sweeps * sum(i[0] for i in insts)
    for _ in range(args.warmup):
        eng.solve_batch_arrays(target, models, plist, configs=False)
    t0 = time.perf_counter()
    dev_ms, tms = [], None
    for _ in range(args.steps):
        costs, cfg, tms = eng.solve_batch_arrays(target, models, plist, configs=True)
        dev_ms.append(tms["device_ms"])
    t = time.perf_counter() - t0
    log("c4 phases ms: init %.1f run %.1f finalize %.1f device %.1f" % (
        tms["init_ms"], tms["run_ms"], tms["finalize_ms"], tms["device_ms"]))
    value = attempts * args.steps / t
    avg_ms = float(np.mean(dev_ms))
    B = 7 * (4 + 1 + 1) + 2   # d = 6 couplings + 1 field term, int8 coefficients
    achieved = attempts * B / (avg_ms * 1e-3) / 1e9
    # plain form: the same solves, one instance per launch, on a sample --
    # with the batch's parameters, and with one replica per work-group (the
    # plain form's best shape; the grouping does not change results)
    stride = max(1, len(models) // args.plain_sample)
    idx = list(range(0, len(models), stride))
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
        plain[label] = plain_att / plain_t
    plain_best = max(plain.values())
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_attempt": B,
                     "kernel": "k_sweep<ISING,pairwise,i32>", "kernel_ms": avg_ms,
                     "lds_frac": achieved / LDS_PEAK_GBS},
        "plain_per_instance": {"value": plain_best, "by_shape": plain, "instances_sampled": len(idx),
                               "speedup_of_batched": value / plain_best,
                               "identical_best_costs": same},
        "autotune": tune,
        "quality": {"best_cost_mean": float(np.mean(costs))},
    }
    return out


def c5_requests(rate, duration, seed=5005, pa_extra=None):
    """Poisson arrivals at `rate` per second over `duration` seconds.  70 %
    Ising ER (N log-uniform in [100, 50000], mean degree 8, J ~ N(0,1) fp32),
    30 % random PUBO (V log-uniform in [500, 20000] vars, 10 V terms of degree
    1/2/3/4 at 10/40/30/20 %, c in [-10, 10] \\ {0}); each request is
    population annealing (population 512, beta 0 -> 3 over 200 steps of 10
    sweeps) or simulated annealing (64 restarts x 1000 sweeps, beta 0.1 ->
    3.0) by a fair coin."""
    from qiotoolkit_amd import workloads as wl
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        k = len(out)
        ising = rng.random() < 0.7
        pa = rng.random() < 0.5
        if ising:
            n = int(round(np.exp(rng.uniform(np.log(100), np.log(50000)))))
            arrays = wl.er_ising_f32_arrays(n, 8.0, seed=seed * 1000 + k)
        else:
            n = int(round(np.exp(rng.uniform(np.log(500), np.log(20000)))))
            arrays = wl.random_pubo_arrays(n, 10 * n, seed=seed * 1000 + k)
        if pa:
            target = "populationannealing.cpu"
            params = {"population": 512, "beta_start": 0.0, "beta_stop": 3.0, "sweeps": 200,
                      "sweeps_per_replica": 10, "seed": 9000 + k, **(pa_extra or {})}
            attempts = 512 * 200 * 10 * n  # nominal: the population fluctuates around 512
        else:
            target = "simulatedannealing.qiotoolkit"
            params = {"beta_start": 0.1, "beta_stop": 3.0, "restarts": 64, "sweeps": 1000, "seed": 9000 + k}
            attempts = 64 * 1000 * n
        out.append({"t": t, "kind": "ising" if ising else "pubo", "n": n, "target": target,
                    "params": params, "arrays": arrays, "nominal_attempts": attempts})
    return out


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
        "failed": fails,
        "server_stats": st,
    }
    return out


def cpu_baseline_c2(inp, temps, args):
    """Reference algorithm (oracle) on a bounded sample: one 32-replica ladder."""
    from oracle import pyoracle as po
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(32, cores))
    p = po.Problem("ising", inp["cost_function"]["terms"], variant="_fast")
    tvals = [temps["initial"] * (temps["final"] / temps["initial"]) ** (i / 31) for i in range(32)]
    # calibrate: a short run, then a sample of ~args.cpu_seconds
    o = po.make_opts(po.RNG_REFERENCE, seed=7, threads=threads, steps_per_tick=1)
    t0 = time.perf_counter()
    p.pt(o, tvals, False, 20)
    dt = max(1e-3, time.perf_counter() - t0)
    sweeps = int(max(20, min(5000, 20 * args.cpu_seconds / dt)))
    t0 = time.perf_counter()
    p.pt(o, tvals, False, sweeps)
    dt = time.perf_counter() - t0
    attempts = 32 * 4096 * sweeps
    # the same ladder on one thread, over a fifth of the sweeps (bounded)
    o1 = po.make_opts(po.RNG_REFERENCE, seed=7, threads=1, steps_per_tick=1)
    sweeps1 = max(20, sweeps // 5)
    t1 = time.perf_counter()
    p.pt(o1, tvals, False, sweeps1)
    dt1 = time.perf_counter() - t1
    return {"value": attempts / dt, "unit": "spin-flip attempts/s", "cores": threads,
            "kind": "port",
            "single_thread": 32 * 4096 * sweeps1 / dt1,
            "sample": "one 32-replica ladder of the same workload, %d of 5000 sweeps, %.1f s, "
                      "pcg32 + ln-table accept, OpenMP over replicas; single_thread: %d sweeps on 1 "
                      "thread, %.1f s" % (sweeps, dt, sweeps1, dt1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--sweeps", type=int, default=5000)
    ap.add_argument("--group", type=int, default=0, help="replicas per work-group (0 = auto)")
    ap.add_argument("--lanes", type=int, default=0, help="positions per class round (0 = auto)")
    ap.add_argument("--positions-per-lane", type=int, default=0,
                    help="lanes cover the largest colour class in this many rounds (0 = auto)")
    ap.add_argument("--instances", type=int, default=4096, help="c4: number of instances")
    ap.add_argument("--no-tune", action="store_true", help="c4: skip the autotuner")
    ap.add_argument("--rate", type=float, default=1.0, help="c5: offered requests per second")
    ap.add_argument("--duration", type=float, default=30.0, help="c5: seconds of arrivals")
    ap.add_argument("--max-active", type=int, default=64, help="c5: concurrently admitted requests")
    ap.add_argument("--sweeps-per-round", type=int, default=10, help="c5: sweeps per shared launch")
    ap.add_argument("--hbm-budget-gib", type=float, default=0.0, help="c5: device memory for requests")
    ap.add_argument("--pa-group", type=int, default=0, help="c5: replicas per work-group of PA requests (0 = solver default)")
    ap.add_argument("--pa-order", default="", help="c5: sweep_order of PA requests (default: the solver's, index)")
    ap.add_argument("--plain-sample", type=int, default=128,
                    help="c4: instances run one per launch for the plain-form rate")
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
    runs = {"c2": run_c2, "c4": run_c4, "c5": run_c5}
    if args.workload not in runs:
        print("unknown workload %s" % args.workload, file=sys.stderr)
        sys.exit(2)
    out = runs[args.workload](args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
