"""Population annealing on the device (qmc_walkers_pa_*) against the oracle's
counter-mode step, step by step: each step's sweeps of the live citizens
equal the oracle's counter-mode sweeps in the same visit order (colour
classes or index order), then the device's population is read back, the oracle computes the step's decisions from those
costs (orc_pa_step_counter: weights, Poisson offspring, copy map, family
statistic, restart rule, global lowest), and the device's population after
its own step must equal the oracle's prediction bit for bit -- states, lowest
states, costs, lowest costs, betas, accepted counters, population size,
restarts and the global lowest state.  Fresh citizens of a restart must be
the oracle's counter-mode initial states with their energies.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from qiotoolkit_amd import _qmc
from qiotoolkit_amd import workloads as wl

from _helpers import arith_of, pair

pytestmark = pytest.mark.gpu


def _steps_match_oracle(ctx, alpha, pop, B, strategy=0, const=0.0, constant_population=False, S=25, side=8,
                        ordered=False, visit=_qmc.VISIT_INDEX_LEVELS, exp_repopulation=False, cap=0,
                        observe=False):
    op, qp = pair(ctx, "ising", wl.lattice2d(side, seed=4), visit=visit)
    cap = cap or max(4 * pop, pop + 256)
    seed = 13
    adaptive = strategy != _qmc.PA_SCHEDULE
    table = np.zeros(2) if adaptive else np.linspace(0.0, 3.0, S + 2)
    w = _qmc.Walkers(qp, cap, B, seed=seed, replicas_per_seed=cap)
    if ordered:
        w.set_sweep_order(True)
    # the sweeps of the live citizens against the oracle's counter-mode
    # sweeps in the same visit order
    sweep_opts = po.make_opts(po.RNG_COUNTER, seed=seed, arith=arith_of(qp),
                              order=None if ordered else qp.order(w.lanes())[0])
    w.pa_begin(pop, alpha, 1.0, table, seed, 1)
    if adaptive or constant_population or exp_repopulation:
        w.pa_strategy(strategy, const, constant_population, exp_repopulation)
    if observe:
        w.pa_observe(S)
    records = []
    w.init(0)
    opts = po.make_opts(po.RNG_COUNTER, seed=seed)
    fam = np.arange(pop, dtype=np.int32)
    R, base, epoch, restarts = pop, 0, 0, 0
    have_best, best, best_state = False, 0.0, None
    cur_beta = table[0]
    target = pop
    evals, pending = 0, pop
    for step in range(S):
        pre = w.read()
        w.sweep(step, None, n_sweeps=1)
        before = w.read()
        want = {k: pre[k][:R].copy() for k in ("states", "cost", "lowest", "lowest_states")}
        op.run_sweeps(sweep_opts, want["states"], want["cost"], want["lowest"], want["lowest_states"],
                      np.arange(R), step, pre["beta"][:R])
        for k, v in want.items():
            assert np.array_equal(before[k][:R], v), (step, k)
        o = po.pa_step_counter(before["cost"][:R], before["lowest"][:R], fam, step, base, table, target, cap,
                               alpha, 1.0, seed, have_best, best, cost_before=pre["cost"][:R], strategy=strategy,
                               strategy_const=const, constant_population=constant_population,
                               exp_repopulation=exp_repopulation)
        w.pa_step()
        info = w.pa_info()
        after = w.read()
        Rf, src = o["R"], o["src"]
        if observe:
            # the step's observations (population_annealing.h:262-470)
            b0, b1 = pre["cost"][:R], before["cost"][:R]
            vb, va = np.mean((b0 - b0.mean()) ** 2), np.mean((b1 - b1.mean()) ** 2)
            gam = np.mean((b0 - b0.mean()) * (b1 - b1.mean()))
            sig = np.sqrt(vb * va)
            tau = zeta = np.nan
            if sig > 0 and gam != 0 and abs(np.log(sig) - np.log(abs(gam))) > 1e-11:
                tau = 1.0 / (np.log(sig) - np.log(abs(gam)))
                zeta = sig * tau
            moved = o["dbeta"] > 0
            nf = fl = 0
            if moved:
                fams = fam[src[:o["R_resampled"]]]
                nf, fl = len(set(fams.tolist())), int(np.sum(fams == fams.max()))
            records.append((epoch, b1.mean(), R, sig, tau, zeta, nf, o["R_resampled"] if moved else None, fl,
                            o["rho"], o["restart"], moved))
        # adaptive steps: device reductions run in another order than the
        # oracle's sequential sums (constant culling lands on the same
        # multiple of 2^-20)
        if strategy == _qmc.PA_CONSTANT_CULLING or not adaptive:
            assert info["delta_beta"] == o["dbeta"], step
        else:
            assert abs(info["delta_beta"] - o["dbeta"]) <= 1e-12 * abs(o["dbeta"]), step
        assert info["population"] == Rf, step
        assert info["target"] == o["target"], step
        # make_step's evaluation count: the sweeps plus the inits before them
        evals += R * op.n + pending
        pending = Rf if o["restart"] else 0
        assert info["evaluations"] == evals, step
        target = o["target"]
        if constant_population and not o["restart"]:
            assert Rf == R
        if o["best_src"] >= 0:
            have_best, best = True, o["best"]
            best_state = before["lowest_states"][o["best_src"]].copy()
        assert info["have_best"] == int(have_best)
        assert info["best"] == best
        assert np.array_equal(info["best_state"], best_state)
        acc = after["accepted"][:Rf]
        assert np.array_equal(acc, before["accepted"][src])
        if o["restart"]:
            epoch += 1
            restarts += 1
            base = step
            cur_beta = table[0]
            for i in range(Rf):
                s = op.initial_state(opts, i, epoch)
                assert np.array_equal(after["states"][i], s), (step, i)
                assert np.array_equal(after["lowest_states"][i], s), (step, i)
                e = op.cost(s)
                assert after["cost"][i] == e and after["lowest"][i] == e, (step, i)
            Rn = o["R_resampled"]
            want_beta = np.where(np.arange(Rf) < Rn, before["beta"][src], table[0])
        else:
            assert np.array_equal(after["states"][:Rf], before["states"][src]), step
            assert np.array_equal(after["lowest_states"][:Rf], before["lowest_states"][src]), step
            assert np.array_equal(after["cost"][:Rf], before["cost"][src]), step
            assert np.array_equal(after["lowest"][:Rf], before["lowest"][src]), step
            cur_beta = cur_beta + info["delta_beta"]
            want_beta = np.full(Rf, cur_beta)
        assert np.array_equal(after["beta"][:Rf], want_beta), step
        assert info["restarts"] == restarts
        assert info["beta"] == cur_beta
        fam, R = o["family"], Rf
    if observe:
        log = w.pa_log(S)
        for s_, (ep, mean, Rs, sig, tau, zeta, nf, rn, fl, rho, rst, moved) in enumerate(records):
            r = log[s_]
            assert r[0] == ep and r[2] == Rs and r[11] == rst and r[12] == moved, s_
            assert abs(r[1] - mean) <= 1e-12 * max(1.0, abs(mean)), s_
            for got, want in ((r[3], sig), (r[4], tau), (r[5], zeta)):
                assert (np.isnan(got) and np.isnan(want)) or abs(got - want) <= 1e-9 * max(1.0, abs(want)), s_
            if moved:
                assert r[7] == nf and r[8] == rn and r[9] == fl and r[10] == rho, s_
    w.pa_end()
    assert w.R == R
    end = w.read()
    assert np.array_equal(end["cost"], after["cost"][:R])
    w.close()
    return restarts


@pytest.mark.parametrize("alpha,pop,B,ordered,visit", [(2.0, 40, 2, False, 1), (1e9, 24, 1, False, 1),
                                                       (2.0, 70, 8, False, 0), (2.0, 40, 32, True, 0),
                                                       (2.0, 70, 16, True, 1)])
def test_device_pa_steps_match_oracle(ctx, alpha, pop, B, ordered, visit):
    """visit 1: index-level classes (the default of the population annealing
    solver), 0: colour classes; ordered: one lane per replica."""
    print("pa steps alpha=%g pop=%d B=%d ordered=%d visit=%d" % (alpha, pop, B, ordered, visit), flush=True)
    _steps_match_oracle(ctx, alpha, pop, B, ordered=ordered, visit=visit)


@pytest.mark.parametrize("strategy,const,constant_population", [
    (_qmc.PA_FRICTION_TENSOR, 1.0, False), (_qmc.PA_ENERGY_VARIANCE, 0.5, False),
    (_qmc.PA_CONSTANT_CULLING, 0.1, False), (_qmc.PA_SCHEDULE, 0.0, True),
    (_qmc.PA_CONSTANT_CULLING, 0.2, True), (_qmc.PA_FRICTION_TENSOR, 1.0, True)])
def test_device_pa_strategies_match_oracle(ctx, strategy, const, constant_population):
    """The adaptive stepping sizes and residual resampling
    (qmc_walkers_pa_strategy) step by step against orc_pa_step_counter, with
    the per-step observation records (qmc_walkers_pa_observe)."""
    print("pa strategy=%d const=%g constant_population=%d" % (strategy, const, constant_population), flush=True)
    _steps_match_oracle(ctx, 2.0, 48, 4, strategy, const, constant_population, observe=True)


def test_device_pa_exponential_repopulation(ctx):
    """Feature switch 2 (FEATURE_PA_EXP_REPOPULATION): every restart doubles
    the target population before the fresh citizens are drawn."""
    print("pa exp repopulation", flush=True)
    restarts = _steps_match_oracle(ctx, 2.0, 16, 4, exp_repopulation=True, cap=4096, S=25)
    assert restarts >= 1


def test_device_pa_constant_population_beyond_the_lds_tree(ctx):
    """Residual resampling of 5000 citizens: the selection tree (8193 nodes)
    lives in device memory instead of LDS."""
    print("pa constant population 5000", flush=True)
    _steps_match_oracle(ctx, 2.0, 5000, 8, _qmc.PA_SCHEDULE, 0.0, True, S=3, side=4)


def test_device_pa_solve_is_deterministic():
    from qiotoolkit_amd.solver import Engine
    e = Engine(0)
    try:
        doc = {"cost_function": {"type": "ising", "terms": wl.lattice2d(10, seed=2)}}
        p = {"params": {"population": 96, "beta_start": 0, "beta_stop": 3, "sweeps": 40, "seed": 5}}
        a = e.solve("populationannealing.cpu", p, doc)
        b = e.solve("populationannealing.cpu", p, doc)
        assert a["solutions"] == b["solutions"]
        assert a["solutions"]["cost"] < -100
    finally:
        e.close()
