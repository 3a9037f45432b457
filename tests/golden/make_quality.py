"""Writes the best-energy distributions of the CPU oracle (reference mode:
pcg32 streams, the 4x256 ln-table accept, index-order sweeps) at the three
single-problem workloads, one best energy per seed for seeds 100..199:

  sa_lattice2d_32   SA, 32x32 periodic +-1 lattice (seed 1), 64 restarts x 1000
                    sweeps, linear beta 0.1 -> 3.0; best over the restarts
  pt_lattice3d_16   PT, 16^3 periodic N(0,1) fp32 spin glass (seed 1), one
                    32-replica ladder at geometric T 0.2 -> 2.0, exchange every
                    sweep, 5000 sweeps; best over the ladder
  sa_pubo_2000      SA, random PUBO 2000 vars / 20000 terms (degrees 1-4 at
                    10/40/30/20 %, c in [-10,10]\\{0}, seed 1), 256 restarts x 2000
                    sweeps, linear beta 0.01 -> 5.0; best over the restarts
  pa_lattice2d_12_* population annealing on the 12x12 periodic +-1 lattice
                    (seed 41) with each stepping size and resampling method
                    (params below); the global lowest of the run
  sa_er_batch_*     SA on three members of the uneven 4096-instance ER batch
                    (smallest, median and largest N), 8 restarts x 1000
                    sweeps, linear beta 0.1 -> 3.0; best over the restarts

The oracle is built with the reference's Release flags (liboracle_fast.so).
The GPU quality test compares the device solver's distribution over the same
seeds with these (two-sample KS p > 0.05, means within 0.5 %).

    python tests/golden/make_quality.py [name ...]     (about an hour on 8 cores)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402
from qiotoolkit_amd import workloads as wl  # noqa: E402

SEEDS = list(range(100, 200))


def cases():
    return {
        "sa_lattice2d_32": dict(
            kind="ising", terms=lambda: wl.lattice2d(32, seed=1), solver="sa",
            params={"beta_start": 0.1, "beta_stop": 3.0, "restarts": 64, "sweeps": 1000}),
        "pt_lattice3d_16": dict(
            kind="ising", terms=lambda: wl.lattice3d(16, seed=1), solver="pt",
            params={"temperatures": {"type": "geometric", "initial": 0.2, "final": 2.0, "count": 32},
                    "sweeps": 5000, "sweeps_per_replica": 1}),
        "sa_pubo_2000": dict(
            kind="pubo", terms=lambda: wl.random_pubo(2000, 20000, seed=1), solver="sa",
            params={"beta_start": 0.01, "beta_stop": 5.0, "restarts": 256, "sweeps": 2000}),
        **pa_cases(),
        **batch_cases(),
    }


# Members of the uneven 4096-instance ER batch (workloads.batched_er_instances):
# the smallest (N = 64), the median (N = 518) and the largest (N = 4086).
BATCH_MEMBERS = {"min": 393, "median": 1918, "max": 4075}


def batch_member(k, seed=4096):
    """Term CSR of member k of the uneven batch, generated alone."""
    N = int(wl.log_uniform_sizes(4096, 64, 4096, seed=seed)[k])
    return wl.er_ising_arrays(N, 6.0, 3, seed=seed * 16 + k)


def batch_cases():
    return {"sa_er_batch_%s" % name: dict(
        kind="ising", terms=(lambda k=k: batch_member(k)), solver="sa", member=k,
        params={"beta_start": 0.1, "beta_stop": 3.0, "restarts": 8, "sweeps": 1000})
        for name, k in BATCH_MEMBERS.items()}


def pa_cases():
    lat = lambda: wl.lattice2d(12, seed=41)  # noqa: E731
    sched = {"beta_start": 0.0, "beta_stop": 3.0}
    out = {
        "pa_lattice2d_12_schedule": dict(population=128, sweeps=200, **sched),
        "pa_lattice2d_12_friction": dict(population=128, sweeps=200, resampling_strategy="friction_tensor",
                                         friction_tensor_constant=1.0),
        "pa_lattice2d_12_energy_variance": dict(population=64, sweeps=100, resampling_strategy="energy_variance",
                                                initial_culling_fraction=1.0),
        "pa_lattice2d_12_culling": dict(population=64, sweeps=200, resampling_strategy="constant_culling",
                                        culling_fraction=0.2),
        "pa_lattice2d_12_schedule_constant": dict(population=32, sweeps=60, constant_population=True, **sched),
        "pa_lattice2d_12_culling_constant": dict(population=32, sweeps=60, resampling_strategy="constant_culling",
                                                 culling_fraction=0.2, constant_population=True),
    }
    return {k: dict(kind="ising", terms=lat, solver="pa", params=v) for k, v in out.items()}


PA_STRATEGIES = {"linear_schedule": (0, None), "friction_tensor": (1, "friction_tensor_constant"),
                 "energy_variance": (2, "initial_culling_fraction"), "constant_culling": (3, "culling_fraction")}


def best_of(op, case, seed, threads):
    p = case["params"]
    if case["solver"] == "sa":
        o = po.make_opts(po.RNG_REFERENCE, seed=seed, threads=threads)
        r = op.sa(o, p["beta_start"], p["beta_stop"], p["sweeps"], p["restarts"])
    elif case["solver"] == "pa":
        strategy, key = PA_STRATEGIES[p.get("resampling_strategy", "linear_schedule")]
        o = po.make_opts(po.RNG_REFERENCE, seed=seed, threads=1)
        r = op.pa(o, p.get("beta_start", 0.0), p.get("beta_stop", 5.0), p["sweeps"], p["population"],
                  constant_population=p.get("constant_population", False), strategy=strategy,
                  strategy_const=p[key] if key else 0.0)
        return float(r["best"] + op.const_cost)
    else:
        t = p["temperatures"]
        temps = wl.geometric_temperatures(t["initial"], t["final"], t["count"])
        o = po.make_opts(po.RNG_REFERENCE, seed=seed, threads=threads, steps_per_tick=1)
        r = op.pt(o, temps, False, p["sweeps"], p["sweeps_per_replica"])
    return float(r["lowest"].min() + op.const_cost)


def main(names):
    threads = os.cpu_count() or 1
    for name, case in cases().items():
        if names and name not in names:
            continue
        op = po.Problem(case["kind"], case["terms"](), variant="_fast")
        t0 = time.time()
        best = []
        for s in SEEDS:
            best.append(best_of(op, case, s, threads))
            print("%s seed %d: %.6f (%.0f s)" % (name, s, best[-1], time.time() - t0), flush=True)
        doc = {"name": name, "kind": case["kind"], "solver": case["solver"], "params": case["params"],
               "seeds": SEEDS, "best": best, "oracle": "reference mode, liboracle_fast.so"}
        with open(os.path.join(HERE, "quality_%s.json" % name), "w") as f:
            json.dump(doc, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
