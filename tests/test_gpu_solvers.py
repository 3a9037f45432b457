"""End-to-end solves through the host library's JSON API on the GPU.

* the reference's solver fixtures (data/*.json with params1.json) reach the
  asserted costs and configurations;
* planted optima shipped with the reference are found;
* every returned configuration's cost equals the oracle's cost function
  (bit-exact for integer couplings, 1e-5 relative for fp32);
* distribution parity: best energies over 100 seeds, GPU vs the oracle running
  the reference algorithm (pcg32, index order), two-sample KS p > 0.05 and
  means within 0.5%.
"""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from qiotoolkit_amd import workloads as wl

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def engine():
    from qiotoolkit_amd.solver import Engine
    e = Engine(0)
    yield e
    e.close()


def load(name):
    return json.load(open(os.path.join(GOLD, name)))


def oracle_cost(doc, configuration):
    """Reference cost function of a rendered configuration (+ const cost)."""
    cf = doc["cost_function"]
    p = po.Problem(cf["type"], cf["terms"])
    names = p.names()
    if cf["type"] == "ising":
        bits = np.array([1 if configuration[str(n)] == -1 else 0 for n in names], dtype=np.uint8)
    else:
        bits = np.array([1 if configuration[str(n)] == 0 else 0 for n in names], dtype=np.uint8)
    return p.cost(bits) + p.const_cost


FIXTURES = [
    ("ising1.json", -10, None),
    ("cpupubojobstest.json", -12, None),
    ("isingemptymixed.json", -100, {"1": -1, "10": 1}),
    ("puboemptymixed.json", -104, {"1": 1, "12": 1}),
    ("isingempty.json", 12, None),
    ("puboempty.json", 2, None),
]


@pytest.mark.parametrize("fname,cost,config", FIXTURES)
@pytest.mark.parametrize("target", ["simulatedannealing.qiotoolkit", "paralleltempering.qiotoolkit",
                                    "populationannealing.cpu"])
def test_reference_fixtures(engine, fname, cost, config, target):
    params = load("params1.json")
    if target == "populationannealing.cpu":
        params = {"params": dict(params["params"], population=64)}
    doc = load(fname)
    res = engine.solve(target, params, doc)
    sol = res["solutions"]
    assert sol["cost"] == cost
    if config:
        for k, v in config.items():
            assert sol["configuration"][k] == v
    if "configuration" in sol:
        assert oracle_cost(doc, sol["configuration"]) == sol["cost"]
        for s in sol["solutions"]:
            assert oracle_cost(doc, s["configuration"]) == s["cost"]


def test_solver_errors(engine):
    from qiotoolkit_amd.solver import SolverError
    doc = load("ising1.json")
    with pytest.raises(SolverError) as e:
        engine.solve("simulatedannealing.qiotoolkit", {"params": {"sweeps": 1}}, doc)
    assert e.value.code == 107  # utils::Error::ValueError
    with pytest.raises(SolverError):
        engine.solve("simulatedannealing.qiotoolkit", {}, doc)


@pytest.mark.parametrize("family", ["wishart_planting_N_16_alpha_0.81_qualitytest",
                                    "combined_4_local_N_16_qualitytest"])
@pytest.mark.parametrize("target", ["simulatedannealing.qiotoolkit", "paralleltempering.qiotoolkit"])
def test_planted_optima(engine, family, target):
    cases = json.load(open(os.path.join(GOLD, "e2e", family, "energies.json")))["cases"]
    for case in cases:
        doc = json.load(open(os.path.join(GOLD, "e2e", family, case["file"])))
        scale = max(abs(t["c"]) for t in doc["cost_function"]["terms"])
        if target.startswith("simulated"):
            p = {"beta_start": 0.1 / scale, "beta_stop": 10.0 / scale, "sweeps": 400, "restarts": 64,
                 "seed": 5}
        else:
            p = {"all_betas": [10.0 / scale * 0.7 ** k for k in range(16)], "sweeps": 400, "seed": 5}
        res = engine.solve(target, {"params": p}, doc)
        assert res["solutions"]["cost"] == case["energy"]
        assert oracle_cost(doc, res["solutions"]["configuration"]) == case["energy"]


def test_fp32_returned_energies_within_tolerance(engine):
    terms = wl.lattice3d(8, seed=3)
    doc = {"cost_function": {"type": "ising", "terms": terms}}
    res = engine.solve("paralleltempering.qiotoolkit",
                       {"params": {"temperatures": {"type": "geometric", "initial": 0.2, "final": 2.0,
                                                    "count": 16}, "sweeps": 300, "seed": 9,
                                   "number_of_solutions": 5}}, doc)
    for s in res["solutions"]["solutions"]:
        ref = oracle_cost(doc, s["configuration"])
        assert abs(ref - s["cost"]) <= 1e-5 * abs(ref)


def ks_2samp(a, b):
    from scipy.stats import ks_2samp as ks
    return ks(a, b).pvalue


@pytest.mark.parametrize("kind", ["lattice2d", "pubo"])
def test_best_energy_distribution_matches_reference(engine, kind):
    if kind == "lattice2d":
        terms = wl.lattice2d(12, seed=41)
        p0 = {"beta_start": 0.1, "beta_stop": 3.0, "sweeps": 60, "restarts": 4}
    else:
        terms = wl.random_pubo(150, 1500, seed=42)
        p0 = {"beta_start": 0.01, "beta_stop": 5.0, "sweeps": 60, "restarts": 4}
    ptype = "ising" if kind == "lattice2d" else "pubo"
    doc = {"cost_function": {"type": ptype, "terms": terms}}
    seeds = list(range(100, 200))
    model = engine.model(doc)
    res, _ = engine.solve_batch("simulatedannealing.qiotoolkit", [model] * len(seeds),
                                [{"params": dict(p0, seed=s)} for s in seeds])
    gpu = np.array([r["solutions"]["cost"] for r in res])
    op = po.Problem(ptype, terms)
    cpu = []
    for s in seeds:
        r = op.sa(po.make_opts(po.RNG_REFERENCE, seed=s), p0["beta_start"], p0["beta_stop"],
                  p0["sweeps"], p0["restarts"])
        cpu.append(r["lowest"].min() + op.const_cost)
    cpu = np.array(cpu)
    p = ks_2samp(gpu, cpu)
    assert p > 0.05, (p, gpu.mean(), cpu.mean())
    assert abs(gpu.mean() - cpu.mean()) <= 0.005 * abs(cpu.mean())


def quality_cases():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_quality", os.path.join(GOLD, "make_quality.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.cases()


@pytest.mark.parametrize("name", ["sa_lattice2d_32", "pt_lattice3d_16", "sa_pubo_2000"])
def test_best_energy_distribution_at_workload_size(engine, name):
    """Full-size workloads (32x32 SA 64 x 1000, 16^3 PT 32 replicas x 5000,
    PUBO 2000-var SA 256 x 2000): best energies over seeds 100..199 against
    the oracle's distribution in tests/golden/quality_<name>.json."""
    path = os.path.join(GOLD, "quality_%s.json" % name)
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated" % os.path.basename(path))
    fx = json.load(open(path))
    case = quality_cases()[name]
    target = {"sa": "simulatedannealing.qiotoolkit", "pt": "paralleltempering.qiotoolkit"}[case["solver"]]
    doc = {"cost_function": {"type": case["kind"], "terms": case["terms"]()}}
    model = engine.model(doc)
    seeds = fx["seeds"]
    print("quality %s: %d solves" % (name, len(seeds)), flush=True)
    res, _ = engine.solve_batch(target, [model] * len(seeds),
                                [{"params": dict(case["params"], seed=s)} for s in seeds])
    gpu = np.array([r["solutions"]["cost"] for r in res])
    cpu = np.array(fx["best"])
    p = ks_2samp(gpu, cpu)
    assert p > 0.05, (p, gpu.mean(), cpu.mean())
    assert abs(gpu.mean() - cpu.mean()) <= 0.005 * abs(cpu.mean())


PA_QUALITY = ["pa_lattice2d_12_schedule", "pa_lattice2d_12_friction", "pa_lattice2d_12_energy_variance",
              "pa_lattice2d_12_culling", "pa_lattice2d_12_schedule_constant", "pa_lattice2d_12_culling_constant"]


@pytest.mark.parametrize("name", PA_QUALITY)
def test_pa_best_energy_distribution_matches_reference(engine, name):
    """Population annealing with each stepping size and resampling method on
    the 12x12 lattice: best energies over seeds 100..199 against the oracle's
    reference-mode runs in tests/golden/quality_<name>.json.  Two-sample KS
    p > 0.05; means within 0.5 %, or within four standard errors of their
    difference where the spread of the runs (restarts) makes 0.5 % smaller
    than the sampling noise."""
    fx = json.load(open(os.path.join(GOLD, "quality_%s.json" % name)))
    case = quality_cases()[name]
    doc = {"cost_function": {"type": case["kind"], "terms": case["terms"]()}}
    model = engine.model(doc)
    print("pa quality %s: %d solves" % (name, len(fx["seeds"])), flush=True)
    gpu = np.array([model.solve("populationannealing.cpu", {"params": dict(case["params"], seed=s)})
                    ["solutions"]["cost"] for s in fx["seeds"]])
    cpu = np.array(fx["best"])
    p = ks_2samp(gpu, cpu)
    assert p > 0.05, (p, gpu.mean(), cpu.mean())
    se = np.sqrt(gpu.var() / len(gpu) + cpu.var() / len(cpu))
    assert abs(gpu.mean() - cpu.mean()) <= max(0.005 * abs(cpu.mean()), 4 * se), (gpu.mean(), cpu.mean(), se)


def test_pa_strategies_find_planted_optima(engine):
    """Every stepping size, with Poisson and constant-population resampling,
    finds the planted optima of the N = 16 quality instances (the device
    counterpart of population_annealing_test.cc:66-152)."""
    variants = [{}, {"resampling_strategy": "friction_tensor"}, {"resampling_strategy": "energy_variance"},
                {"resampling_strategy": "constant_culling", "culling_fraction": 0.2},
                {"constant_population": True},
                {"resampling_strategy": "constant_culling", "culling_fraction": 0.2, "constant_population": True},
                # feature switch 2 as population_annealing_test.cc:66-101 enable it
                {"resampling_strategy": "friction_tensor", "enabled_features": [2]},
                {"resampling_strategy": "energy_variance", "enabled_features": [2]}]
    for family in ("wishart_planting_N_16_alpha_0.81_qualitytest", "combined_4_local_N_16_qualitytest"):
        cases = json.load(open(os.path.join(GOLD, "e2e", family, "energies.json")))["cases"][:2]
        for case in cases:
            doc = json.load(open(os.path.join(GOLD, "e2e", family, case["file"])))
            for v in variants:
                p = dict({"population": 64, "sweeps": 100, "seed": 42, "beta_start": 0.0, "beta_stop": 5.0}, **v)
                res = engine.solve("populationannealing.cpu", {"params": p}, doc)
                assert res["solutions"]["cost"] == case["energy"], (family, case["file"], v)
                assert oracle_cost(doc, res["solutions"]["configuration"]) == case["energy"]
                assert res["solutions"]["parameters"]["constant_population"] == bool(v.get("constant_population"))


@pytest.mark.parametrize("target", ["simulatedannealing.qiotoolkit", "paralleltempering.qiotoolkit"])
def test_batch_arrays_match_json_and_single_solves(engine, target):
    """Batched solves of uneven instances (arrays and JSON results) equal the
    same solves run one per launch; every returned configuration has the
    returned cost under the oracle's cost function."""
    insts = []
    for k, n in enumerate([64, 300, 1000, 90, 2000, 150]):
        off, ids, c = wl.er_ising_arrays(n, 6.0, 3, seed=300 + k)
        insts.append((off, ids, c))
    models = [engine.model_from_arrays("ising", *t) for t in insts]
    if target.startswith("simulated"):
        base = {"beta_start": 0.1, "beta_stop": 3.0, "restarts": 8, "sweeps": 60}
    else:
        base = {"all_betas": [3.0, 2.0, 1.0, 0.5], "sweeps": 60}
    plist = [{"params": dict(base, seed=11 + k, positions_per_lane=2 if k % 2 else 0)}
             for k in range(len(models))]
    cost, cfg, tm = engine.solve_batch_arrays(target, models, plist)
    res, _ = engine.solve_batch(target, models, plist)
    for k, m in enumerate(models):
        single = m.solve(target, plist[k])
        assert single["solutions"]["cost"] == cost[k] == res[k]["solutions"]["cost"]
        names = m.names()
        want = {str(n): int(v) for n, v in zip(names, cfg[k])}
        assert want == res[k]["solutions"]["configuration"]
        off, ids, c = insts[k]
        terms = [{"c": float(c[j]), "ids": [int(x) for x in ids[off[j]:off[j + 1]]]} for j in range(len(c))]
        doc = {"cost_function": {"type": "ising", "terms": terms}}
        assert oracle_cost(doc, want) == cost[k]
    assert tm["device_ms"] > 0


def test_solve_from_protobuf_folder(engine):
    """The reference's ring fixture (data/models_input_problem2_pb) solved
    from its protobuf folder and from the equivalent JSON document."""
    folder = os.path.join(GOLD, "proto", "models_input_problem2_pb")
    c = [-9.5 if i == 9 else (-(i + 1.0) if i % 2 == 0 else i + 1.0) for i in range(10)]
    doc = {"cost_function": {"type": "ising", "version": "1.0",
                             "terms": [{"c": c[i], "ids": [i, (i + 1) % 10]} for i in range(10)]}}
    p = {"params": {"beta_start": 0.1, "beta_stop": 3.0, "sweeps": 50, "restarts": 8, "seed": 3}}
    for target in ("simulatedannealing.qiotoolkit", "populationannealing.cpu"):
        a = engine.solve_file(target, p, folder)
        b = engine.solve(target, p, doc)
        assert a["solutions"] == b["solutions"]
        assert oracle_cost(doc, a["solutions"]["configuration"]) == a["solutions"]["cost"]


def test_benchmark_solver_properties(engine):
    """benchmark.extensions "solver" as SteppingSolver::get_solver_properties
    renders it (stepping_solver.h:268-283, solver.h:57-75), in the shape the
    reference's SA toy test prints (simulated_annealing_test.cc:159-177, a
    release build: no milestones recorded)."""
    doc = load("ising1.json")
    p = {"params": {"step_limit": 100, "restarts": 4, "seed": 42, "beta_start": 0.5, "cost_limit": -1000,
                    "beta_stop": 10.0}}
    res = engine.solve("simulatedannealing.qiotoolkit", p, doc)
    ext = {e["name"]: e["value"] for e in res["benchmark"]["extensions"]}
    sol = ext["solver"]
    assert sol["cost_limit"] == -1000
    assert sol["cost_milestones"] == []
    assert sol["last_step"] == 100 and sol["step_limit"] == 100
    assert sol["exit_reason"] == "Stop due to step limit: 100, steps:100"
    # reaching the cost limit ends the run early with the reference's message
    p["params"]["cost_limit"] = 0
    res = engine.solve("simulatedannealing.qiotoolkit", p, doc)
    sol = {e["name"]: e["value"] for e in res["benchmark"]["extensions"]}["solver"]
    assert sol["exit_reason"] == "Stop due to cost limit: 0.000000, current lowest cost:"
    assert sol["last_step"] < 100


def test_pa_observer_render(engine):
    """A population annealing solve with the reference's observer protocol
    (observe/observer.h:80-92): step windows of 10, averages of the step
    observations, constants of the last step and epoch of each window, in
    benchmark.extensions "solver"."""
    doc = {"cost_function": {"type": "ising", "terms": wl.lattice2d(10, seed=2)}}
    p = {"params": {"population": 64, "beta_start": 0, "beta_stop": 3, "sweeps": 40, "seed": 5},
         "observer": {"average_over": {"step": 10}, "constant": ["step", "epoch"],
                      "average": ["cost", "population", "delta_beta", "rho_t", "families", "family_size",
                                  "sigma"]}}
    res = engine.solve("populationannealing.cpu", p, doc)
    sol = {e["name"]: e["value"] for e in res["benchmark"]["extensions"]}["solver"]
    groups = sol["step"]
    assert len(groups) >= 4
    assert groups[-1]["step"] == 39 and sol["last_step"] == 40
    for g in groups:
        assert 32 <= g["population"] <= 128
        assert g["sigma"] >= 0
        assert g["epoch"] >= 0
    # delta_beta is 3/39 per step of the schedule wherever it was observed
    seen = [g["delta_beta"] for g in groups if "delta_beta" in g]
    assert seen and all(abs(d - 3.0 / 39) < 1e-12 for d in seen)


def test_eval_limit_stops_at_the_tick(engine):
    """eval_limit (stepping_solver.h:235-244): checked after each tick of
    10, 12, 14, 16, ... steps against cost evaluations (one per replica at
    init) plus difference evaluations (replicas x variables per step).
    ising1 has 3 variables: SA with 4 restarts counts 4 + 12 s, so a limit of
    500 stops after the tick ending at 52 steps (628 evaluations); PT with
    the default 5 betas counts 5 + 15 s and stops at 36 (545)."""
    doc = load("ising1.json")
    res = engine.solve("simulatedannealing.qiotoolkit",
                       {"params": {"restarts": 4, "seed": 3, "sweeps": 1000, "eval_limit": 500}}, doc)
    sol = {e["name"]: e["value"] for e in res["benchmark"]["extensions"]}["solver"]
    assert sol["last_step"] == 52 and sol["eval_limit"] == 500
    assert sol["exit_reason"] == "Stop due to Evaluation limit: 500, evaluations:628"
    res = engine.solve("paralleltempering.qiotoolkit", {"params": {"seed": 3, "sweeps": 1000, "eval_limit": 500}},
                       doc)
    sol = {e["name"]: e["value"] for e in res["benchmark"]["extensions"]}["solver"]
    assert sol["last_step"] == 36
    assert sol["exit_reason"] == "Stop due to Evaluation limit: 500, evaluations:545"
    assert res["solutions"]["cost"] == -10


def test_pa_eval_limit(engine):
    """Population annealing stops at the first tick end (10, 22, 36, 52, ...
    steps) where its evaluations reach eval_limit."""
    doc = {"cost_function": {"type": "ising", "terms": wl.lattice2d(6, seed=2)}}
    p = {"params": {"population": 16, "beta_start": 0, "beta_stop": 3, "sweeps": 500, "seed": 5,
                    "eval_limit": 20000}}
    res = engine.solve("populationannealing.cpu", p, doc)
    sol = {e["name"]: e["value"] for e in res["benchmark"]["extensions"]}["solver"]
    ends, s, tick = [], 0, 10
    while s < 500:
        s += tick
        ends.append(s)
        tick = int(max(tick + 1.0, tick * 1.2))
    assert sol["last_step"] in ends and sol["last_step"] < 500
    reason = sol["exit_reason"]
    assert reason.startswith("Stop due to Evaluation limit: 20000, evaluations:")
    assert int(reason.rsplit(":", 1)[1]) >= 20000


def test_pt_observer_render(engine):
    """The reference's PT observer configuration
    (parallel_tempering_test.cc:113-131) on a device solve: one entry per
    replica with T, acc_rate, avg_cost, avg_dir and (all but the last)
    swap_rate; the lowest temperature's direction is always up (0), the
    highest's average is near 1, rates lie in [0, 1]."""
    doc = {"cost_function": {"type": "ising", "terms": wl.lattice2d(8, seed=3)}}
    p = {"params": {"seed": 42, "step_limit": 200,
                    "temperatures": {"type": "linear", "initial": 3.0, "final": 0.5, "count": 5}},
         "observer": {"group_by": ["replica"], "constant": ["T"],
                      "average": ["avg_cost", "acc_rate", "avg_dir", "swap_rate"]}}
    res = engine.solve("paralleltempering.qiotoolkit", p, doc)
    sol = {e["name"]: e["value"] for e in res["benchmark"]["extensions"]}["solver"]
    reps = sol["replica"]
    assert [r["T"] for r in reps] == [0.5, 1.125, 1.75, 2.375, 3.0]
    assert all(set(r) == {"T", "acc_rate", "avg_cost", "avg_dir", "swap_rate"} for r in reps[:-1])
    assert "swap_rate" not in reps[-1]
    assert reps[0]["avg_dir"] == 0 and reps[-1]["avg_dir"] > 0.9
    assert all(0 <= r["acc_rate"] <= 1 for r in reps)
    assert reps[0]["avg_cost"] < reps[-1]["avg_cost"]
    assert res["solutions"]["cost"] <= reps[0]["avg_cost"]


PF_TARGETS = ["simulatedannealing-parameterfree.qiotoolkit", "paralleltempering-parameterfree.qiotoolkit"]


@pytest.mark.gpu
@pytest.mark.parametrize("fname,cost,config", FIXTURES)
@pytest.mark.parametrize("target", PF_TARGETS)
def test_parameter_free_reference_fixtures(engine, fname, cost, config, target):
    """sa_pf_test.cc / pt_pf_test.cc: the reference's fixtures with pf5s.json
    ({"timeout": 5}) reach the same costs and configurations."""
    doc = load(fname)
    res = engine.solve(target, {"params": {"timeout": 5, "seed": 11}}, doc)
    sol = res["solutions"]
    assert sol["cost"] == cost
    for k, v in (config or {}).items():
        assert sol["configuration"][k] == v
    if "configuration" in sol:
        assert oracle_cost(doc, sol["configuration"]) == sol["cost"]
        for s in sol["solutions"]:
            assert oracle_cost(doc, s["configuration"]) == s["cost"]


@pytest.mark.gpu
@pytest.mark.parametrize("target", PF_TARGETS)
def test_parameter_free_search(engine, target):
    """The search stops after 8 runs without a lower cost, well inside the
    time limit; the best run's parameters are reported; the planted optimum
    of a 16-variable Wishart instance is found."""
    family = "wishart_planting_N_16_alpha_0.81_qualitytest"
    case = json.load(open(os.path.join(GOLD, "e2e", family, "energies.json")))["cases"][0]
    doc = json.load(open(os.path.join(GOLD, "e2e", family, case["file"])))
    res = engine.solve(target, {"params": {"timeout": 30, "seed": 3}}, doc)
    assert res["solutions"]["cost"] == case["energy"]
    sol = [e for e in res["benchmark"]["extensions"] if e["name"] == "solver"][0]["value"]
    tr = sol["training"]
    assert tr["runs"] >= 8 and tr["search_ms"] < 30e3
    assert "without a lower cost" in sol["search_exit_reason"]
    par = res["solutions"]["parameters"]
    assert par["sweeps"] >= 2
    if target.startswith("simulated"):
        assert par["restarts"] >= 64 and par["beta_stop"] > par["beta_start"] > 0
    else:
        assert len(par["all_betas"]) == 32
    # same seed, same search
    again = engine.solve(target, {"params": {"timeout": 30, "seed": 3}}, doc)
    assert again["solutions"]["parameters"] == par
    assert [e for e in again["benchmark"]["extensions"] if e["name"] == "solver"][0]["value"]["training"]["runs"] \
        == tr["runs"]


@pytest.mark.gpu
def test_parameter_free_time_limit(engine):
    """A spent time limit ends the search after its first pass
    (parameter_free_solver.h:101-108): the four runs of [10, 20, 30, 50]."""
    doc = load("ising1.json")
    res = engine.solve("simulatedannealing-parameterfree.qiotoolkit", {"params": {"timeout": 1e-9, "seed": 2}}, doc)
    assert res["solutions"]["cost"] == -10
    sol = [e for e in res["benchmark"]["extensions"] if e["name"] == "solver"][0]["value"]
    assert sol["training"]["runs"] == 4 and sol["training"]["passes"] == 1
    assert sol["search_exit_reason"].startswith("Stop due to time limit")


def test_pa_time_and_cost_limits(engine):
    """Population annealing checks time_limit and cost_limit after each tick
    of 10, 12, 14, 16, ... steps (stepping_solver.h:151-152, 195-258): a spent
    time limit stops after the first tick, a cost limit above the reachable
    costs stops there too, each with the reference's exit reason."""
    doc = {"cost_function": {"type": "ising", "terms": wl.lattice2d(8, seed=2)}}
    base = {"population": 32, "beta_start": 0, "beta_stop": 3, "sweeps": 400, "seed": 5}
    res = engine.solve("populationannealing.cpu", {"params": dict(base, timeout=1e-9)}, doc)
    sol = {e["name"]: e["value"] for e in res["benchmark"]["extensions"]}["solver"]
    assert sol["last_step"] == 10
    assert sol["exit_reason"].startswith("Stop due to time_limit.steps : 10, time_limit:")
    res = engine.solve("populationannealing.cpu", {"params": dict(base, cost_limit=1000)}, doc)
    sol = {e["name"]: e["value"] for e in res["benchmark"]["extensions"]}["solver"]
    assert sol["last_step"] == 10
    assert sol["exit_reason"].startswith("Stop due to cost limit: 1000.000000")
    assert res["solutions"]["cost"] <= 1000
    # without limits the run takes all its steps
    res = engine.solve("populationannealing.cpu", {"params": base}, doc)
    sol = {e["name"]: e["value"] for e in res["benchmark"]["extensions"]}["solver"]
    assert sol["last_step"] == 400


@pytest.mark.gpu
def test_parameter_free_predicted_overrun():
    """On the 16^3 lattice longer runs keep finding lower costs, so the
    search stretches its sweeps every pass; with a 1 s limit it stops before
    a pass that, at the previous pass's rate, would not end within the limit
    -- and so never overruns it."""
    from qiotoolkit_amd.solver import Engine
    eng = Engine(0)
    doc = {"cost_function": {"type": "ising", "terms": wl.lattice3d(16, seed=1)}}
    res = eng.solve("simulatedannealing-parameterfree.qiotoolkit", {"params": {"timeout": 1.0, "seed": 4}}, doc)
    eng.close()
    sol = [e for e in res["benchmark"]["extensions"] if e["name"] == "solver"][0]["value"]
    assert sol["search_exit_reason"] == ("Stop due to time limit: 1.000000 "
                                         "(the next pass would not end within it)"), sol["search_exit_reason"]
    assert sol["training"]["passes"] >= 2
    assert sol["training"]["search_ms"] < 1000.0


def test_benchmark_block_of_the_runner(engine):
    """The benchmark block as Runner::get_run_output and Solver::get_benchmark
    render it (app/runner.cc:541-628, solver/solver.h:78-146, keys of
    utils/config.h): graph size properties (puboemptymixed: 4 variables, a
    constant term -> min_locality 0), timing, CPU time, memory, input and
    response bytes, threads from the params, build and invocation
    extensions; and the _QIO_METRICS line of the response."""
    from qiotoolkit_amd import solver
    doc = load("puboemptymixed.json")
    terms = doc["cost_function"]["terms"]
    params = {"params": {"beta_start": 0.1, "beta_stop": 3.0, "sweeps": 50, "restarts": 8, "seed": 3,
                         "threads": 4}}
    res = engine.solve("simulatedannealing.qiotoolkit", params, doc)
    b = res["benchmark"]
    nonconst = [t for t in terms if t["ids"]]
    names = {i for t in terms for i in t["ids"]}
    assert b["num_variables"] == len(names) and b["num_terms"] == len(nonconst)
    assert b["locality"] == max(len(t["ids"]) for t in terms)
    assert b["min_locality"] == min(len(t["ids"]) for t in terms)
    assert b["sum_coefficient_degrees_total"] == sum(len(t["ids"]) for t in terms)
    assert b["accumulated_dependent_vars"] == sum(len(t["ids"]) * (len(t["ids"]) - 1) for t in nonconst)
    assert b["avg_locality"] == sum(len(t["ids"]) for t in terms) / len(nonconst)
    assert b["max_coupling_magnitude"] == max(abs(t["c"]) for t in nonconst)
    assert b["min_coupling_magnitude"] == min(abs(t["c"]) for t in nonconst)
    assert b["threads"] == 4
    assert b["solver_time_ms"] > 0 and b["cpu_time_ms"] > 0 and b["max_mem_bytes"] > 1 << 20
    assert b["end2end_time_ms"] >= b["solver_time_ms"]
    assert b["end2end_cputime_ms"] >= 0 and b["preprocessing_time_ms"] >= 0 and b["postprocessing_time_ms"] >= 0
    assert b["disk_io_read_bytes"] == len(json.dumps(params)) + len(json.dumps(doc))
    assert b["disk_io_write_bytes"] > 200
    ext = {e["name"]: e["value"] for e in b["extensions"]}
    assert {"counters", "solver", "model", "build", "invocation", "device"} <= set(ext)
    assert ext["model"]["graph"]["size"]["nodes"] == b["num_variables"]
    assert ext["build"]["build_type"] == "Release" and ext["invocation"]["host"]
    line = solver.metrics(res)
    m = json.loads(line[len("_QIO_METRICS"):-len("_QIO_METRICS")])
    assert m["problem"]["model"] == "pubo" and m["problem"]["num_variables"] == b["num_variables"]
    assert m["common"]["solver_time_ms"] == b["solver_time_ms"]
    assert m["solver"]["last_step"] == 50
