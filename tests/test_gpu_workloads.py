"""The two batching/serving workloads at their full shapes on the GPU.

* The uneven batch: all 4096 Erdos-Renyi instances (N log-uniform in
  [64, 4096], mean degree 6, J in {-3..3}\\{0}, h in {-1,0,1}), 8 restarts x
  1000 sweeps of simulated annealing each (linear beta 0.1 -> 3.0), in the
  launch shape the autotuner picks: every returned configuration's energy
  under the oracle's cost function equals the returned cost bit for bit; a
  sample of the instances solved one per launch returns the same costs and
  configurations; and on three members (smallest, median, largest N) the
  best energies over seeds 100..199 match the oracle's reference-mode runs
  (simulated_annealing.h:76-110) -- two-sample KS p > 0.05, means within
  0.5 %.
* The serving stream's request shapes: population annealing (population 512,
  200 steps x 10 sweeps, beta 0 -> 3) and simulated annealing (64 x 1000) on
  a 50 000-variable fp32 ER Ising model and a 20 000-variable PUBO with
  200 000 terms, submitted together to the serving loop: each result equals
  the same request solved alone (PUBO bit-exact, fp32 within 1e-5 relative),
  and each returned configuration's cost under the oracle's cost function
  is the returned cost (PUBO exactly, fp32 within 1e-5 relative).
"""
import importlib.util
import json
import os

import numpy as np
import pytest
from scipy.stats import ks_2samp

from oracle import pyoracle as po
from qiotoolkit_amd import workloads as wl

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
SA = "simulatedannealing.qiotoolkit"
PA = "populationannealing.cpu"


def _make_quality():
    spec = importlib.util.spec_from_file_location("make_quality", os.path.join(GOLD, "make_quality.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _oracle_cost(op, names, config, kind):
    """Cost of a configuration given by internal variable id (+1/-1 Ising,
    1/0 PUBO) under the oracle's cost function, constant terms included."""
    pos = {int(n): i for i, n in enumerate(op.names())}
    bits = np.zeros(op.n, dtype=np.uint8)
    down = -1 if kind == "ising" else 0
    for n, v in zip(names, config):
        bits[pos[int(n)]] = 1 if int(v) == down else 0
    return op.cost(bits) + op.const_cost


@pytest.fixture(scope="module")
def engine():
    from qiotoolkit_amd.solver import Engine
    e = Engine(0)
    yield e
    e.close()


def test_uneven_batch_full_size(engine):
    insts = wl.batched_er_instances(4096)
    print("batch: %d instances, %d spins" % (len(insts), sum(i[0] for i in insts)), flush=True)
    models = [engine.model_from_arrays("ising", off, ids, c) for (_, off, ids, c) in insts]

    def params(k, extra):
        return {"params": dict({"beta_start": 0.1, "beta_stop": 3.0, "restarts": 8, "sweeps": 1000,
                                "seed": 7000 + k}, **extra)}

    tune = engine.tune_batch(SA, models, [params(k, {}) for k in range(len(models))], trial_sweeps=20)
    ch = tune["params"]
    extra = {"replicas_per_group": ch["replicas_per_group"], "positions_per_lane": ch["positions_per_lane"]}
    if ch["sweeps_per_launch_fraction"] > 0:
        extra["sweeps_per_launch"] = max(1, int(round(ch["sweeps_per_launch_fraction"] * 1000)))
    print("tuned shape %s" % json.dumps(ch), flush=True)
    plist = [params(k, extra) for k in range(len(models))]
    if ch.get("large_min_vars", 0) > 0:  # the largest solves in smaller work-groups
        for k, inst in enumerate(insts):
            if inst[0] >= ch["large_min_vars"]:
                plist[k]["params"]["replicas_per_group"] = ch["large_replicas_per_group"]
                plist[k]["params"]["positions_per_lane"] = ch["large_positions_per_lane"]
    cost, cfg, tm = engine.solve_batch_arrays(SA, models, plist, configs=True)
    print("batch solved: device %.1f ms" % tm["device_ms"], flush=True)
    # every returned configuration under the oracle's cost function
    for k, (_, off, ids, c) in enumerate(insts):
        op = po.Problem("ising", (off, ids, c))
        assert _oracle_cost(op, models[k].names(), cfg[k], "ising") == cost[k], k
    # one instance per launch: same costs and configurations
    for k in range(0, len(models), 256):
        pk = {"params": dict(plist[k]["params"])}
        pk["params"].pop("sweeps_per_launch", None)
        c1, g1, _ = engine.solve_batch_arrays(SA, [models[k]], [pk], configs=True)
        assert c1[0] == cost[k], k
        assert np.array_equal(g1[0], cfg[k]), k
    for m in models:
        m.close()


@pytest.mark.parametrize("member", ["min", "median", "max"])
def test_uneven_batch_member_distribution(engine, member):
    mq = _make_quality()
    name = "sa_er_batch_%s" % member
    fx = json.load(open(os.path.join(GOLD, "quality_%s.json" % name)))
    case = mq.cases()[name]
    k = case["member"]
    off, ids, c = case["terms"]()
    # the fixture's instance is member k of the batch the test above solves
    _, boff, bids, bc = wl.batched_er_instances(4096)[k]
    assert np.array_equal(off, boff) and np.array_equal(ids, bids) and np.array_equal(c, bc)
    model = engine.model_from_arrays("ising", off, ids, c)
    seeds = fx["seeds"]
    res, _ = engine.solve_batch(SA, [model] * len(seeds),
                                [{"params": dict(case["params"], seed=s)} for s in seeds])
    gpu = np.array([r["solutions"]["cost"] for r in res])
    cpu = np.array(fx["best"])
    p = ks_2samp(gpu, cpu).pvalue
    print("%s: gpu mean %.3f cpu mean %.3f KS p %.3f" % (name, gpu.mean(), cpu.mean(), p), flush=True)
    assert p > 0.05, (p, gpu.mean(), cpu.mean())
    assert abs(gpu.mean() - cpu.mean()) <= 0.005 * abs(cpu.mean())
    model.close()


def _stream_shaped_requests():
    """(kind, arrays, target, params) at the serving stream's largest shapes."""
    ising = wl.er_ising_f32_arrays(50000, 8.0, seed=5101)
    pubo = wl.random_pubo_arrays(20000, 200000, seed=5102)
    out = []
    for j, (kind, arrays, n) in enumerate((("ising", ising, 50000), ("pubo", pubo, 20000))):
        for pa in (True, False):
            r = wl.stream_request(pa, n, 600 + 2 * j + int(pa))
            out.append((kind, arrays, r["target"], r["params"]))
    return out


def test_serving_at_stream_shapes():
    from qiotoolkit_amd.solver import Engine, Server
    reqs = _stream_shaped_requests()
    eng = Engine(0)
    alone = []
    for kind, arrays, target, params in reqs:
        m = eng.model_from_arrays(kind, *arrays)
        alone.append((m.names(), m.solve(target, {"params": params})))
        print("alone %s %s: cost %.6f" % (kind, target, alone[-1][1]["solutions"]["cost"]), flush=True)
        m.close()
    eng.close()
    srv = Server(0, max_active=8, sweeps_per_round=10)
    ids = [srv.submit(target, {"params": params}, kind, *arrays) for kind, arrays, target, params in reqs]
    mixed = []
    for rid in ids:
        rc, doc = srv.wait(rid, 300000)
        assert rc == 0, doc
        mixed.append(doc)
    st = srv.stats()
    srv.close()
    print("mix: p50 %.0f ms p99 %.0f ms" % (st["latency_p50_ms"], st["latency_p99_ms"]), flush=True)
    for (kind, arrays, target, params), (names, a), b in zip(reqs, alone, mixed):
        ca, cb = a["solutions"]["cost"], b["solutions"]["cost"]
        conf = b["solutions"]["configuration"]
        op = po.Problem(kind, arrays)
        ref = _oracle_cost(op, names, [conf[str(int(n))] for n in names], kind)
        if kind == "pubo":
            assert ca == cb, (target, ca, cb)
            assert a["solutions"]["configuration"] == conf
            assert ref == cb
        else:
            assert abs(ca - cb) <= 1e-5 * abs(ca), (target, ca, cb)
            assert abs(ref - cb) <= 1e-5 * abs(ref), (target, ref, cb)
        assert b["benchmark"]["serving"]["finished_ms"] >= b["benchmark"]["serving"]["admitted_ms"]
