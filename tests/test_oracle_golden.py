"""Pin the CPU oracle before trusting it (runs on CPU).

Each check is independent of the oracle's own formulas: reference-held golden
vectors (utils/test/random_generator_test.cc), published known-answer vectors
(Random123 Philox), the reference's model/schedule/solver test answers, brute
force with the textbook energy formula, and planted optima shipped with the
reference (data/input_e2e_pipeline/*/energies.json).
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)["cost_function"]


# ---------------------------------------------------------------- RNG
def pct(xs):
    return [int(x * 100) for x in xs]


def test_pcg_default_seed():  # random_generator_test.cc:38-44
    assert pct(po.pcg_uniform(3)) == [9, 83, 5]


def test_pcg_custom_seed():  # random_generator_test.cc:46-57
    assert pct(po.pcg_uniform(3, 42)) == [41, 26, 40]


def test_pcg_fork():  # random_generator_test.cc:59-78
    f = po.pcg_fork_uniform(42, 3, 12)
    assert pct(f[0]) == [95, 79, 48, 70, 2, 89, 70, 10, 27, 66, 47, 18]
    assert pct(f[1]) == [86, 95, 8, 65, 36, 92, 17, 14, 82, 9, 90, 19]
    assert pct(f[2]) == [79, 27, 7, 13, 39, 85, 26, 66, 89, 5, 55, 13]
    assert pct(f[3]) == [14, 82, 1, 6, 90, 27, 67, 82, 94, 24, 76, 35]


def test_philox_known_answers():  # Random123 kat_vectors, philox4x32_10
    assert list(po.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(po.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E,
                                                                   0xA20BC7C6, 0x6D5451FD]
    assert list(po.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                          [0xA4093822, 0x299F31D0])) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                         0x24126EA1]


def test_poisson_moments():
    # parity unpinned: libstdc++ std::poisson_distribution used as the reference
    # uses it; only its distribution is checked here.
    import ctypes as C
    out = np.zeros(20000, dtype=np.uint32)
    po.lib().orc_pcg_poisson(7, 3.5, len(out), out.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert abs(out.mean() - 3.5) < 0.06 and abs(out.var() - 3.5) < 0.15


# ---------------------------------------------------------------- LUT accept
def test_uniform_ln_closed_form():  # utils/utils.cc:12-32
    t = po.uniform_ln()
    for r in range(4):
        assert t[r, 0] == -math.inf
        for i in (1, 2, 100, 255):
            assert t[r, i] == math.log(i / 256.0 ** (r + 1))


def test_accept_rules():  # markov/metropolis.h:64-104
    assert po.accept(0.0, 1.0, []) == (True, 0)          # downhill: no draw
    assert po.accept(-3.0, 1.0, []) == (True, 0)
    assert po.accept(1.0, 1.0, [1]) == (False, 1)        # bucket 1: ln(0) -> NaN -> reject
    assert po.accept(1.0, 1.0, [0, 0, 0, 0]) == (True, 4)  # four zero buckets accept
    assert po.accept(1e9, 1.0, [0, 0, 0, 0]) == (True, 4)
    assert po.accept(1.0, 0.0, [200]) == (True, 1)       # beta 0 accepts any uphill move
    assert po.accept(1.0, math.inf, [200]) == (False, 1)  # T = 0 rejects uphill


def test_accept_exhaustive_cells():
    """Every (row, bucket) cell and both ends of the interpolation, against the closed form."""
    t = po.uniform_ln()
    D = float(0xFFFFFE)
    for row in range(4):
        for bucket in range(1, 256):
            for hi in (0, 0x7FFFFF, 0xFFFFFF):
                r = (hi << 8) | bucket
                y0, y1 = t[row, bucket - 1], t[row, bucket]
                with np.errstate(invalid="ignore"):
                    value = y0 + (hi / D) * (y1 - y0)
                for bde in (1e-3, 0.5, 3.0, 5.5, 12.0):
                    draws = [0] * row + [r]
                    acc, used = po.accept(bde, 1.0, draws)
                    assert used == row + 1
                    assert acc == bool(value < -bde)


def test_accept_probability_approximates_boltzmann():
    rng = np.random.default_rng(5)
    words = rng.integers(0, 2**32, size=(40000, 4), dtype=np.uint64).astype(np.uint32)
    for bde in (0.1, 1.0, 3.0):
        p = np.mean([po.accept(bde, 1.0, w)[0] for w in words])
        assert abs(p - math.exp(-bde)) < 0.01


# ---------------------------------------------------------------- models
def ring(n, c=1):
    return [{"c": c, "ids": [i, (i + 1) % n]} for i in range(n)]


def test_ising_ring10_known_answers():  # ising_term_cached_test.cc:61-76
    p = po.Problem("ising", ring(10))
    s = np.zeros(10, dtype=np.uint8)
    assert p.cost(s) == 10
    assert p.delta(s, 0) == -4
    s[0] = 1
    assert p.cost(s) == 6
    assert p.delta(s, 0) == 4


def test_pubo_ring10_known_answers():  # pubo_with_counter_test.cc:59-74
    p = po.Problem("pubo", ring(10))
    s = np.zeros(10, dtype=np.uint8)
    assert p.cost(s) == 10
    assert p.delta(s, 0) == -2
    s[0] = 1
    assert p.cost(s) == 8
    assert p.delta(s, 0) == 2


def test_single_spin_terms():  # ising_test.cc:572-618
    p = po.Problem("ising", [{"c": 1, "ids": [i]} for i in range(7)])
    s = np.ones(7, dtype=np.uint8)
    assert p.cost(s) == -7
    for i in range(7):
        assert p.delta(s, i) == 2
        s[i] = 0
        assert p.cost(s) == (i + 1) * 2 - 7


def test_constant_terms_and_names():  # ising_test.cc:284-360, pubo_with_counter_test.cc:206-244
    p = po.Problem("ising", [{"c": 1, "ids": [-1, 2]}, {"c": 2, "ids": [2, 100]}, {"c": 200, "ids": []}])
    assert list(p.names()) == [-1, 2, 100]
    assert p.const_cost == 200
    p = po.Problem("pubo", [{"c": 1, "ids": []}, {"c": 2, "ids": []}])
    assert p.n == 0 and p.const_cost == 3


def test_duplicate_ids():  # graph.h:131-138; pubo.h:82 merges
    with pytest.raises(ValueError, match="Duplicated"):
        po.Problem("ising", [{"c": 1, "ids": [0, 0]}])
    p = po.Problem("pubo", [{"c": 3, "ids": [4, 4, 5]}])
    off, ids, c = p.edges()
    assert list(ids) == [0, 1] and p.n == 2


def test_empty_terms_rejected():  # graph.h:88-92
    with pytest.raises(ValueError, match="greater than 0"):
        po.Problem("ising", [])


def textbook_energy(kind, terms, names, bits):
    """H = sum_j c_j prod_{i in j} s_i, s = +1/-1 (Ising) or x = 1/0 (PUBO)."""
    pos = {n: k for k, n in enumerate(names)}
    e = 0.0
    for t in terms:
        if kind == "ising":
            prod = 1
            for i in set(t["ids"]):
                prod *= -1 if bits[pos[i]] else 1
        else:
            prod = 1
            for i in set(t["ids"]):
                prod *= 0 if bits[pos[i]] else 1
        e += t["c"] * prod
    return e


@pytest.mark.parametrize("kind", ["ising", "pubo"])
def test_brute_force_cost_and_delta(kind):
    rng = np.random.default_rng(3 if kind == "ising" else 4)
    terms = []
    for k in range(40):
        size = int(rng.integers(1, 5))
        ids = rng.choice(10, size=size, replace=False).tolist()
        terms.append({"c": int(rng.integers(-7, 8)), "ids": ids})
    terms.append({"c": 5, "ids": []})
    p = po.Problem(kind, terms)
    names = p.names()
    n = p.n
    for bits in itertools.product([0, 1], repeat=n):
        b = np.array(bits, dtype=np.uint8)
        want = textbook_energy(kind, terms, names, b) - p.const_cost
        got = p.cost(b)
        assert got == want
        i = int(rng.integers(0, n))
        b2 = b.copy()
        b2[i] ^= 1
        assert p.delta(b, i) == textbook_energy(kind, terms, names, b2) - textbook_energy(kind, terms, names, b)


def test_f32_delta_is_fp32_rounding_of_f64():
    rng = np.random.default_rng(9)
    terms = [{"c": float(np.float32(rng.standard_normal())), "ids": [int(a), int(b)]}
             for a, b in rng.integers(0, 30, size=(120, 2)) if a != b]
    p = po.Problem("ising", terms)
    s = rng.integers(0, 2, size=p.n, dtype=np.uint8)
    for i in range(p.n):
        d64 = p.delta(s, i, po.ARITH_F64)
        d32 = p.delta(s, i, po.ARITH_F32)
        assert abs(d32 - d64) <= 1e-5 * max(1.0, abs(d64))


# ---------------------------------------------------------------- schedules
def test_linear_schedule_known_values():  # schedule/test/linear_test.cc:16-61
    v = lambda x: po.lib().orc_schedule_value(0, 3, 5, -1, -1, x)
    assert [v(-1), v(0), v(0.5), v(1), v(2)] == [3, 3, 4, 5, 5]
    v = lambda x: po.lib().orc_schedule_value(0, 5, 3, 2, 4, x)
    assert [v(1), v(2), v(3), v(4), v(5)] == [5, 5, 4, 3, 3]


def test_geometric_schedule_known_values():  # schedule/test/geometric_test.cc:16-124
    v = lambda x: po.lib().orc_schedule_value(1, 3, 6, -1, -1, x)
    assert v(0) == 3 and v(1) == 6 and abs(v(0.5) - 3 * 2 ** 0.5) < 1e-15
    import ctypes as C
    out = np.zeros(100)
    po.lib().orc_schedule_discretized(1, 10, 1, 100, out.ctypes.data_as(C.POINTER(C.c_double)))
    for k, want in [(1, 9.7700995729922528), (5, 8.9021508544503831), (12, 7.5646332755462842),
                    (24, 5.7223676593502102), (37, 4.2292428743894899), (45, 3.5111917342151227),
                    (52, 2.9836472402833309), (67, 2.1049041445120129), (70, 1.9630406500402637),
                    (82, 1.4849682622544584), (95, 1.0974987654930506), (99, 0.9999999999999947)]:
        assert abs(out[k] - want) < 1e-14


# ---------------------------------------------------------------- solvers
FIXTURES = [  # (file, kind, expected cost, expected configuration subset)
    ("ising1.json", -10, None),
    ("cpupubojobstest.json", -12, None),
    ("isingemptymixed.json", -100, {1: -1, 10: 1}),
    ("puboemptymixed.json", -104, {1: 1, 12: 1}),
    ("isingempty.json", 12, None),
    ("puboempty.json", 2, None),
]


def render(kind, names, bits):
    if kind == "ising":
        return {int(n): (-1 if b else 1) for n, b in zip(names, bits)}
    return {int(n): (0 if b else 1) for n, b in zip(names, bits)}


@pytest.mark.parametrize("fname,cost,config", FIXTURES)
@pytest.mark.parametrize("solver", ["sa", "pt", "pa"])
def test_solver_fixtures(fname, cost, config, solver):
    # simulated_annealing_test.cc:194-260, parallel_tempering_test.cc:235-281,
    # population_annealing_test.cc:243-329 with data/params1.json
    params = json.load(open(os.path.join(GOLD, "params1.json")))["params"]
    cf = load(fname)
    p = po.Problem(cf["type"], cf["terms"])
    o = po.make_opts(po.RNG_REFERENCE, seed=params["seed"], threads=1)
    if solver == "sa":
        r = p.sa(o, params["beta_start"], params["beta_stop"], params["sweeps"], params["restarts"])
        k = int(np.argmin(r["lowest"])) if p.n else 0
        best, state = (r["lowest"][k] if p.n else 0.0), r["lowest_states"][k]
    elif solver == "pt":
        r = p.pt(o, params["all_betas"], True, params["sweeps"])
        k = int(np.argmin(r["lowest"])) if p.n else 0
        best, state = (r["lowest"][k] if p.n else 0.0), r["lowest_states"][k]
    else:
        r = p.pa(o, params["beta_start"], params["beta_stop"], params["sweeps"], 16)
        best, state = (r["best"] if p.n else 0.0), r["best_state"]
    assert best + p.const_cost == cost
    if config:
        got = render(cf["type"], p.names(), state)
        for key, val in config.items():
            assert got[key] == val


@pytest.mark.parametrize("family", ["wishart_planting_N_16_alpha_0.81_qualitytest",
                                    "combined_4_local_N_16_qualitytest"])
def test_planted_optima(family):
    cases = json.load(open(os.path.join(GOLD, "e2e", family, "energies.json")))["cases"]
    for case in cases:
        cf = json.load(open(os.path.join(GOLD, "e2e", family, case["file"])))["cost_function"]
        p = po.Problem(cf["type"], cf["terms"])
        o = po.make_opts(po.RNG_REFERENCE, seed=22)
        scale = max(abs(t["c"]) for t in cf["terms"])
        r = p.sa(o, 0.1 / scale, 10.0 / scale, 300, 32)
        assert r["lowest"].min() + p.const_cost == case["energy"]
        # the optimum is a true minimum: exhaustive check on N = 16
        best = np.inf
        for bits in itertools.product([0, 1], repeat=p.n):
            best = min(best, p.cost(np.array(bits, dtype=np.uint8)))
            if best + p.const_cost < case["energy"]:
                raise AssertionError("planted energy is not minimal")
        assert best + p.const_cost == case["energy"]


def test_counter_mode_initial_state_layout():
    """Counter-mode initial state: bit i%32 of Philox INIT word i/32 (model/ising.h:23-41)."""
    p = po.Problem("ising", ring(70))
    o = po.make_opts(po.RNG_COUNTER, seed=99, stream=3)
    s = p.initial_state(o, replica=5, epoch=2)
    for i in range(70):
        wi = i // 32
        w = po.philox([wi >> 2, 5, 2, 2], [99, 3])[wi & 3]
        assert s[i] == ((int(w) >> (i % 32)) & 1)


def test_steps_per_tick_minimum():
    # stepping_solver.h:112-268: the first tick always runs steps_per_tick steps
    p = po.Problem("ising", ring(6))
    r = p.sa(po.make_opts(seed=1), 0.1, 1.0, 3, 2)
    assert r["steps"] == 10
    r = p.sa(po.make_opts(seed=1), 0.1, 1.0, 37, 2)
    assert r["steps"] == 37


# ------------------------------------------------- counter-mode population step


def _canonical2(w0, w1):
    return min((float(w0) + float(w1) * 4294967296.0) / 18446744073709551616.0, 1.0 - 2.0 ** -53)


@pytest.mark.parametrize("lam", [0.05, 0.7, 1.0, 3.3, 12.0, 29.5])
def test_poisson_counter_is_inverse_cdf(lam):
    """Below 30 the counter-mode sampler is the inverse Poisson CDF of its
    first uniform: the smallest k with F(k) >= u (scipy's ppf), checked
    draw by draw away from the CDF steps."""
    from scipy.stats import poisson
    seed, step, n = 17, 5, 400
    got = po.poisson_counter(lam, seed, step, n)
    checked = 0
    for i in range(n):
        w = po.philox([i, step, 0, 4], [seed, 0])
        u = _canonical2(w[0], w[1])
        k = int(poisson.ppf(u, lam))
        if abs(poisson.cdf(k, lam) - u) > 1e-12 and (k == 0 or abs(poisson.cdf(k - 1, lam) - u) > 1e-12):
            assert got[i] == k, (i, u, got[i], k)
            checked += 1
    assert checked > n - 5


@pytest.mark.parametrize("lam", [30.0, 45.5, 200.0, 5000.0])
def test_poisson_counter_ptrs_matches_pmf(lam):
    """From 30 up (transformed rejection) the draws follow the Poisson pmf:
    chi-square over bins with >= 50 expected counts."""
    from scipy.stats import chisquare, poisson
    n = 200000
    x = po.poisson_counter(lam, 23, 9, n).astype(np.int64)
    lo, hi = int(poisson.ppf(1e-4, lam)), int(poisson.ppf(1 - 1e-4, lam))
    edges = [lo]
    while edges[-1] < hi:
        k = edges[-1] + 1
        while k < hi and n * (poisson.cdf(k - 1, lam) - poisson.cdf(edges[-1] - 1, lam)) < 50:
            k += 1
        edges.append(k)
    obs, exp = [], []
    prev_cdf = 0.0
    bounds = [-1] + edges[1:-1] + [10 ** 9]
    for a, b in zip(bounds[:-1], bounds[1:]):
        obs.append(int(((x > a) & (x <= b)).sum()))
        exp.append(n * (poisson.cdf(b, lam) - poisson.cdf(a, lam)))
    exp = np.array(exp) * (sum(obs) / sum(exp))
    assert chisquare(obs, exp).pvalue > 1e-4


def _pa_inputs(R, seed=3, spread=4.0):
    rng = np.random.default_rng(seed)
    cost = np.round(rng.normal(0, spread, R)) - 50
    lowest = cost - np.abs(np.round(rng.normal(0, 2, R)))
    fam = rng.integers(0, R, R).astype(np.int32)
    return cost, lowest, fam


def test_pa_step_counter_offspring_follow_the_boltzmann_weights():
    """Mean offspring of citizen i over many steps = target * w_i / sum(w),
    w_i = exp(-dbeta (E_i - E_min)) (population_annealing.h:490-616)."""
    R, target = 40, 40
    cost, lowest, fam = _pa_inputs(R)
    table = np.linspace(0.0, 2.0, 2001)  # dbeta = 0.001 at every step
    counts = np.zeros(R)
    n = 1500
    for step in range(n):
        r = po.pa_step_counter(cost, lowest, fam, step, step, table, target, 4 * target, 1e-9, 1.0, 11,
                               True, -1e9)
        assert not r["restart"] and r["moved"]
        counts += np.bincount(r["src"], minlength=R)
        assert np.all(np.diff(r["src"]) >= 0)  # offspring in citizen order
        assert np.array_equal(r["family"], fam[r["src"]])
    w = np.exp(-0.001 * (cost - cost.min()))
    want = target * w / w.sum()
    assert np.allclose(counts / n, want, atol=4 * np.sqrt(want / n).max())


def test_pa_step_counter_family_statistic_and_restart_rule():
    R, target = 30, 30
    cost, lowest, _ = _pa_inputs(R, seed=5)
    table = np.linspace(0.0, 3.0, 11)
    fam = np.zeros(R, dtype=np.int32)  # one family: rho = R'^2 / R
    r = po.pa_step_counter(cost, lowest, fam, 2, 0, table, target, 200, 1.5, 1.0, 7, False, 0.0)
    assert r["rho"] == r["R_resampled"] ** 2 / R
    assert r["restart"] == (R / r["rho"] < 1.5)
    if r["restart"]:
        assert list(r["family"][:target]) == list(range(min(target, r["R"])))
        assert r["R"] >= target
    # zero cost variance restarts; a restart tops the population up to the
    # target with copies of the last resampled citizen and resets families
    flat = np.full(R, -7.0)
    r = po.pa_step_counter(flat, flat, np.arange(R, dtype=np.int32), 4, 0, np.ones(6), 64, 256, 2.0, 1.0, 7,
                           True, -7.0)
    assert r["restart"] and r["R"] == 64 and r["moved"] and r["R_resampled"] == R
    assert list(r["src"][:R]) == list(range(R)) and np.all(r["src"][R:] == R - 1)
    assert list(r["family"]) == list(range(64))
    assert r["best_src"] == -1 and r["best"] == -7.0  # not lower than the best so far


def test_pa_step_counter_identity_without_beta_change_and_global_lowest():
    R = 20
    cost, lowest, fam = _pa_inputs(R, seed=8)
    table = np.array([1.0, 1.0, 1.0])
    r = po.pa_step_counter(cost, lowest, fam, 0, 0, table, R, 100, 1.5, 1.0, 3, True, -1e9)
    assert not r["moved"] and not r["restart"] and r["rho"] == 1.0
    assert list(r["src"]) == list(range(R)) and np.array_equal(r["family"], fam)
    assert r["best_src"] == -1
    r = po.pa_step_counter(cost, lowest, fam, 0, 0, table, R, 100, 1.5, 1.0, 3, False, 0.0)
    k = int(np.argmin(lowest))
    assert r["best_src"] == k and r["best"] == lowest[k]


def test_pa_step_counter_keeps_an_emptied_population_alive():
    """All offspring counts zero: one copy of the citizen picked by Philox
    (0, step, 0, 5) (population_annealing.h:566-572)."""
    cost = np.array([-3.0])
    table = np.linspace(0, 1, 4001)
    seen = 0
    for step in range(200):
        r = po.pa_step_counter(cost, cost, np.zeros(1, np.int32), step, step, table, 1, 16, 1e-9, 1.0, 5,
                               True, -10.0)
        cnt = po.poisson_counter(1.0, 5, step, 1)[0]
        if cnt == 0:
            seen += 1
            assert r["R"] == 1 and list(r["src"]) == [0]
        else:
            assert r["R"] == cnt
    assert seen > 40  # P(0) = 1/e


def _selector_by_scan(weights, uniforms):
    """utils::RandomSelector's draws by a linear scan: ids kept in insertion
    slots, a removed slot refilled by the last one; the pick is the last slot
    whose prefix weight is <= u * total."""
    ids = list(range(len(weights)))
    w = list(weights)
    out = []
    for u in uniforms:
        x = u * sum(w)
        pos = max(p for p in range(len(w)) if sum(w[:p]) <= x)
        out.append(ids[pos])
        ids[pos], w[pos] = ids[-1], w[-1]
        ids.pop()
        w.pop()
    return out


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (5, 2), (16, 3), (37, 4), (300, 5)])
def test_selector_matches_a_linear_scan(n, seed):
    """The selection tree (random_selector.cc) against a linear scan, with
    dyadic weights so every sum is exact: same picks, including zero weights
    and removals of the last slot."""
    rng = np.random.default_rng(seed)
    weights = rng.integers(0, 64, n) / 64.0
    weights[rng.random(n) < 0.2] = 0.0
    weights[0] = max(weights[0], 1 / 64)
    k = int(np.count_nonzero(weights))
    uniforms = rng.integers(0, 2 ** 20, k) / 2.0 ** 20
    got = po.selector_draw(weights, uniforms)
    assert list(got) == _selector_by_scan(weights, uniforms)
    assert len(set(got)) == k and all(weights[i] > 0 for i in got)


def test_selector_single_pick_is_the_inverse_cdf():
    """One pick with u on a fine grid: id i is returned for exactly the
    fraction w_i / sum(w) of the grid."""
    w = np.array([3.0, 0.0, 1.0, 4.0, 0.5, 7.5])
    M = 1 << 12
    hits = np.zeros(len(w))
    for j in range(M):
        hits[po.selector_draw(w, [(j + 0.5) / M])[0]] += 1
    assert np.array_equal(hits / M, w / w.sum())


@pytest.mark.parametrize("eps0,E,scale", [(0.1, 1.0, 1.0), (0.2, 3.0, 0.5), (0.05, 12.0, 2.0), (0.3, 0.25, 1.0)])
def test_culling_estimate_and_find_delta_beta_closed_form(eps0, E, scale):
    """Half the population at cost 0, half at E, population = target: the
    culling estimate is tanh(x/2)/2 - eps0 with x = dbeta E scale, so the root
    is 2 atanh(2 eps0) / (E scale); find_delta_beta returns the lower end of
    a 2^-20 bracket of it (population_annealing.h:175-260)."""
    R = 64
    shifted = np.array([0.0] * (R // 2) + [E] * (R // 2))
    for db in (0.0, 0.01, 0.3, 2.0):
        x = db * E * scale
        assert abs(po.culling_estimate(shifted, eps0, R, scale, db) - (np.tanh(x / 2) / 2 - eps0)) < 1e-14
    root = 2 * np.arctanh(2 * eps0) / (E * scale)
    before = np.zeros(R)
    got = po.adaptive_delta_beta(3, eps0, before, shifted - 5.0, R, scale)  # unshifted costs
    assert got <= root < got + 2.0 ** -20 + 1e-12
    assert (got * 2 ** 20) == int(got * 2 ** 20)  # a multiple of 2^-20


def test_friction_tensor_and_energy_variance_stepping():
    """Costs before (0,1,2,3) and after (0,2,1,3) the sweeps: variances 1.25,
    covariance 1, sigma 1.25, so zeta = 1.25 s / ln 1.25 for s sweeps
    (population_annealing.h:347-373, 420-439)."""
    b = np.array([0.0, 1.0, 2.0, 3.0])
    a = np.array([0.0, 2.0, 1.0, 3.0])
    for s in (1, 10):
        ft = po.adaptive_delta_beta(1, 0.7, b, a, 4, 1.0, s)
        assert abs(ft - 0.7 * np.sqrt(np.log(1.25) / (1.25 * s))) < 1e-15
    assert abs(po.adaptive_delta_beta(2, 0.5, b, a, 4) - 0.5 / np.sqrt(1.25)) < 1e-15
    # sigma = |gamma| (delta = 0), or no correlation (gamma = 0): no step
    assert po.adaptive_delta_beta(1, 1.0, b, 2 * b + 1, 4) == 0.0
    assert po.adaptive_delta_beta(1, 1.0, b, np.array([1.0, 0.0, 0.0, 1.0]), 4) == 0.0
    # zero variance: no step
    assert po.adaptive_delta_beta(2, 0.5, b, np.full(4, 3.0), 4) == 0.0


def test_pa_step_counter_constant_population():
    """Residual resampling (population_annealing.h:540-578): every count is
    floor(target w / sum w) or one more, the counts sum to the population
    size, and the fill picks follow the selection tree with uniforms from
    Philox (j, step, 1, 5)."""
    R = target = 48
    cost, lowest, fam = _pa_inputs(R, seed=12)
    table = np.linspace(0.0, 0.2, 40)
    for step in range(30):
        r = po.pa_step_counter(cost, lowest, fam, step, 0, table, target, 4 * target, 1e-9, 1.0, 21, True, -1e9,
                               constant_population=True)
        assert r["R_resampled"] == R and not r["restart"]
        cnt = np.bincount(r["src"], minlength=R)
        db = table[step + 1] - table[step]
        w = np.exp(-db * (cost - cost.min()))
        rw = target * w / w.sum()
        fl = np.floor(rw)
        assert np.all((cnt == fl) | (cnt == fl + 1))
        fills = R - int(fl.sum())
        extra = np.flatnonzero(cnt == fl + 1)
        assert len(extra) == fills
        u = [float(pcanon(po.philox([j, step, 1, 5], [21, 0]))) for j in range(fills)]
        assert sorted(po.selector_draw(rw - fl, u)) == sorted(extra)


def pcanon(words):
    return (int(words[0]) + int(words[1]) * 2.0 ** 32) / 2.0 ** 64


def test_pa_step_counter_adaptive_strategies_use_their_stepping():
    R = target = 40
    cost, lowest, fam = _pa_inputs(R, seed=3)
    before = cost + np.round(np.random.default_rng(4).normal(0, 3, R))
    z = np.zeros(2)
    for strategy, c in ((1, 1.0), (2, 0.5), (3, 0.1)):
        r = po.pa_step_counter(cost, lowest, fam, 5, 0, z, target, 4 * target, 1e-9, 1.0, 9, True, -1e9,
                               cost_before=before, strategy=strategy, strategy_const=c, sweeps_per_step=3)
        assert r["dbeta"] == po.adaptive_delta_beta(strategy, c, before, cost, target, 1.0, 3)
        assert r["dbeta"] > 0 and r["moved"]


@pytest.mark.parametrize("strategy,const", [(0, 0.0), (1, 1.0), (2, 0.5), (3, 0.2)])
@pytest.mark.parametrize("constant_population", [False, True])
def test_pa_strategies_reach_planted_optima(strategy, const, constant_population):
    """Population annealing with every stepping size, with Poisson or
    constant-population resampling, finds the planted (exhaustively checked)
    optima of the N = 16 quality instances -- the counterpart of
    population_annealing_test.cc:66-152 on the reference's toy model."""
    for family in ("wishart_planting_N_16_alpha_0.81_qualitytest", "combined_4_local_N_16_qualitytest"):
        cases = json.load(open(os.path.join(GOLD, "e2e", family, "energies.json")))["cases"][:2]
        for case in cases:
            cf = json.load(open(os.path.join(GOLD, "e2e", family, case["file"])))["cost_function"]
            p = po.Problem(cf["type"], cf["terms"])
            o = po.make_opts(po.RNG_REFERENCE, seed=42, threads=1)
            r = p.pa(o, 0.0, 5.0, 100, 64, constant_population=constant_population, strategy=strategy,
                     strategy_const=const)
            assert r["best"] + p.const_cost == case["energy"], (family, case["file"])
            if constant_population:
                assert r["final_population"] == 64


def test_toy_model_sa_milestones():
    """The reference's SA on its toy model (simulated_annealing_test.cc:143-180:
    seed 42, 4 restarts, 100 steps, beta 0.5 -> 10, cost_limit 0): the
    lowest-cost milestones it prints.  Pins pcg32 seeding and fork, the
    two-draw uniform, the ln-table accept on continuous cost differences, the
    stretched linear schedule and the strict lowest tracking."""
    want = [(1, 9.464013), (2, 0.794104), (3, 0.245903), (4, 0.132852), (6, 0.097936), (7, 0.000760),
            (84, 0.000191)]
    got = po.toy_sa(42, 100, 4, 0.5, 10.0)
    assert [int(s) for s, _ in got] == [s for s, _ in want]
    for (_, c), (_, w) in zip(got, want):
        assert abs(c - w) <= 5e-7


def test_toy_model_pt_replica_statistics():
    """The reference's PT on its toy model (parallel_tempering_test.cc:113-190:
    seed 42, 200 steps, temperatures linear 0.5 -> 0.1, 5 replicas): the
    per-replica step averages its observer reports and the lowest-cost
    milestones.  Pins the replica forks, the exchange rule
    u < exp((E_i - E_i+1)(beta_i - beta_i+1)) on pairs i = step mod 2 with the
    master's uniforms, the direction labels and the averaging."""
    stats, ms = po.toy_pt(42, 200, np.linspace(0.5, 0.1, 5))
    want = {"T": [0.1, 0.2, 0.3, 0.4, 0.5],
            "acc_rate": [0.053, 0.0765, 0.11, 0.159, 0.208],
            "avg_cost": [0.334950, 0.448586, 0.426404, 0.460988, 0.596713],
            "avg_dir": [0.0, 0.515, 0.565, 0.705, 0.995],
            "swap_rate": [0.63, 0.85, 0.80, 0.88]}
    for k, v in want.items():
        assert np.allclose(stats[k][:len(v)], v, rtol=0, atol=5e-7), (k, stats[k])
    want_ms = [(1, 9.464013), (2, 0.145333), (3, 0.076090), (5, 0.020307), (12, 0.003722), (88, 0.000020)]
    assert [int(s) for s, _ in ms] == [s for s, _ in want_ms]
    for (_, c), (_, w) in zip(ms, want_ms):
        assert abs(c - w) <= 5e-7


def test_pa_exponential_repopulation():
    """Feature switch 2 (population_annealing.h:463-470): each restart
    doubles the target, so the final population is the initial one times a
    power of two at least 2^restarts / (population fluctuation), and the
    counter-mode step reports the doubled target."""
    fam = os.path.join(GOLD, "e2e", "wishart_planting_N_16_alpha_0.81_qualitytest")
    first = json.load(open(os.path.join(fam, "energies.json")))["cases"][0]["file"]
    cf = json.load(open(os.path.join(fam, first)))["cost_function"]
    p = po.Problem(cf["type"], cf["terms"])
    o = po.make_opts(po.RNG_REFERENCE, seed=42, threads=1)
    r = p.pa(o, 0.0, 5.0, 60, 4, exp_repopulation=True)
    plain = p.pa(po.make_opts(po.RNG_REFERENCE, seed=42, threads=1), 0.0, 5.0, 60, 4)
    assert r["restarts"] >= 1
    assert r["final_population"] >= 4 * 2 ** r["restarts"] // 2
    assert r["final_population"] > plain["final_population"] or r["restarts"] == 0
    flat = np.full(6, -3.0)
    st = po.pa_step_counter(flat, flat, np.arange(6, dtype=np.int32), 2, 0, np.linspace(0, 1, 5), 6, 64, 2.0, 1.0,
                            3, True, -3.0, exp_repopulation=True)
    assert st["restart"] and st["target"] == 12 and st["R"] == 12
    assert list(st["family"]) == list(range(12))
