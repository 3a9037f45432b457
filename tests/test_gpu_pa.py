"""Population annealing on the device (qmc_walkers_pa_*) against the oracle's
counter-mode step, step by step: after each step's sweeps the device's
population is read back, the oracle computes the step's decisions from those
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

from _helpers import pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alpha,pop,B", [(2.0, 40, 2), (1e9, 24, 1), (2.0, 70, 8)])
def test_device_pa_steps_match_oracle(ctx, alpha, pop, B):
    print("pa steps alpha=%g pop=%d B=%d" % (alpha, pop, B), flush=True)
    op, qp = pair(ctx, "ising", wl.lattice2d(8, seed=4))
    cap = max(4 * pop, pop + 256)
    seed, S = 13, 25
    table = np.linspace(0.0, 3.0, S + 2)
    w = _qmc.Walkers(qp, cap, B, seed=seed, replicas_per_seed=cap)
    w.pa_begin(pop, alpha, 1.0, table, seed, 1)
    w.init(0)
    opts = po.make_opts(po.RNG_COUNTER, seed=seed)
    fam = np.arange(pop, dtype=np.int32)
    R, base, epoch, restarts = pop, 0, 0, 0
    have_best, best, best_state = False, 0.0, None
    cur_beta = table[0]
    for step in range(S):
        w.sweep(step, None, n_sweeps=1)
        before = w.read()
        o = po.pa_step_counter(before["cost"][:R], before["lowest"][:R], fam, step, base, table, pop, cap,
                               alpha, 1.0, seed, have_best, best)
        w.pa_step()
        info = w.pa_info()
        after = w.read()
        Rf, src = o["R"], o["src"]
        assert info["population"] == Rf, step
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
            cur_beta = cur_beta + o["dbeta"]
            want_beta = np.full(Rf, cur_beta)
        assert np.array_equal(after["beta"][:Rf], want_beta), step
        assert info["restarts"] == restarts
        assert info["beta"] == cur_beta
        fam, R = o["family"], Rf
    w.pa_end()
    assert w.R == R
    end = w.read()
    assert np.array_equal(end["cost"], after["cost"][:R])
    w.close()


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
