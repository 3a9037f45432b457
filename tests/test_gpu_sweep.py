"""GPU parity: the HIP sweep engine against the CPU oracle, through the C ABI.

Integer couplings: bit-exact states, costs, lowest costs and lowest states
against the oracle run in the same visit order with the same counter-based
RNG contract.  FP32 couplings: identical trajectories (the oracle sums the
fp32 coefficients in the same order) and energies within 1e-5 relative.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from qiotoolkit_amd import _qmc, workloads as wl
from tests._helpers import arith_of, oracle_initial, pair

pytestmark = pytest.mark.gpu

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def test_philox_device_matches_kat_and_oracle(ctx):
    for c, k, want in KAT:
        got = ctx.philox([c], k)[0]
        assert list(got) == want
    rng = np.random.default_rng(1)
    ctrs = rng.integers(0, 2**32, size=(512, 4), dtype=np.uint64).astype(np.uint32)
    key = [123456789, 987654321]
    got = ctx.philox(ctrs, key)
    want = np.stack([po.philox(c, key) for c in ctrs])
    assert np.array_equal(got, want)


def _klocal(n, m, kmax, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(m):
        ids = sorted(rng.choice(n, size=1 + k % kmax, replace=False).tolist())
        c = int(rng.integers(1, 8)) * (1 if rng.random() < 0.5 else -1)
        out.append(dict(c=c, ids=ids))
    return out


PROBLEMS = {
    "lattice2d_pm1": ("ising", lambda: wl.lattice2d(8, seed=3)),
    "er_int": ("ising", lambda: wl.er_ising(150, 6.0, seed=4)),
    "pubo_int": ("pubo", lambda: wl.random_pubo(120, 600, seed=5)),
    "ising_4local": ("ising", lambda: [dict(c=int(c), ids=ids) for c, ids in zip(
        np.random.default_rng(6).integers(-5, 6, 300),
        [sorted(np.random.default_rng(100 + k).choice(40, size=1 + k % 4, replace=False).tolist())
         for k in range(300)]) if c != 0]),
    "lattice3d_f32": ("ising", lambda: wl.lattice3d(6, seed=7)),
    # terms of up to 6 variables: the incidence-CSR path (more than 3 other
    # members per term do not fit the 16-byte sliced entries)
    "pubo_6local": ("pubo", lambda: _klocal(60, 240, 6, seed=12)),
    "ising_6local": ("ising", lambda: _klocal(50, 200, 6, seed=13)),
}


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_energy_and_delta_match_oracle(ctx, name):
    kind, gen = PROBLEMS[name]
    op, qp = pair(ctx, kind, gen())
    rng = np.random.default_rng(11)
    states = rng.integers(0, 2, size=(16, op.n), dtype=np.uint8)
    e_dev = qp.energy(states)
    e_ref = np.array([op.cost(s) for s in states])
    if qp.info.coupling == _qmc.QMC_COUPLING_I32:
        assert np.array_equal(e_dev, e_ref)
    else:
        np.testing.assert_allclose(e_dev, e_ref, rtol=1e-12, atol=1e-9)
    d_dev = qp.delta_all(states)
    ar = arith_of(qp)
    d_ref = np.array([[op.delta(s, i, ar) for i in range(op.n)] for s in states])
    assert np.array_equal(d_dev, d_ref)
    # the GPU's dE agrees with the reference's FP64 arithmetic within 1e-5 relative
    d64 = np.array([[op.delta(s, i, po.ARITH_F64) for i in range(op.n)] for s in states])
    np.testing.assert_allclose(d_dev, d64, rtol=1e-5, atol=1e-5 * np.abs(d64).max())


def _run_pair(ctx, kind, terms, R, B, betas, seed=7, ordered=False, lanes=0, visit=0):
    op, qp = pair(ctx, kind, terms, visit=visit)
    w = _qmc.Walkers(qp, R, B, seed=seed)
    if lanes:
        w.set_lanes(lanes)
    order, co = qp.order(w.lanes())
    assert op.check_classes(order, co)
    w.init()
    r0 = w.read()
    opts = po.make_opts(po.RNG_COUNTER, seed=seed, arith=arith_of(qp),
                        order=None if ordered else order)
    init = oracle_initial(op, opts, R)
    assert np.array_equal(r0["states"], init)
    costs = np.array([op.cost(s) for s in init])
    if qp.info.coupling == _qmc.QMC_COUPLING_I32:
        assert np.array_equal(r0["cost"], costs)
    states = init.copy()
    lowest = costs.copy()
    lstates = init.copy()
    w.sweep(0, schedule=betas, ordered=ordered)
    r1 = w.read()
    acc = op.run_sweeps(opts, states, costs, lowest, lstates, np.arange(R), 0,
                        np.repeat(np.asarray(betas, dtype=np.float64), R))
    return op, qp, w, r1, dict(states=states, cost=costs, lowest=lowest, lowest_states=lstates,
                               accepted=acc)


@pytest.mark.parametrize("name", ["lattice2d_pm1", "er_int", "pubo_int", "ising_4local", "pubo_6local",
                                  "ising_6local"])
@pytest.mark.parametrize("B", [1, 5, 8, 32])
def test_sa_sweeps_bit_exact_int(ctx, name, B):
    kind, gen = PROBLEMS[name]
    betas = np.linspace(0.1, 3.0, 24)
    op, qp, w, got, want = _run_pair(ctx, kind, gen(), R=37, B=B, betas=betas)
    assert qp.info.coupling == _qmc.QMC_COUPLING_I32
    assert np.array_equal(got["states"], want["states"])
    assert np.array_equal(got["cost"], want["cost"])
    assert np.array_equal(got["lowest"], want["lowest"])
    assert np.array_equal(got["lowest_states"], want["lowest_states"])
    assert np.array_equal(got["accepted"], want["accepted"])
    # every returned configuration's energy equals the oracle cost function
    assert np.array_equal(w.energy(lowest=True), np.array([op.cost(s) for s in got["lowest_states"]]))


@pytest.mark.parametrize("name", ["lattice2d_pm1", "er_int", "pubo_int", "ising_4local", "pubo_6local"])
@pytest.mark.parametrize("B", [1, 8])
def test_index_levels_make_the_index_order_sweep(ctx, name, B):
    """Classes = dependency levels of the index order (qmc_problem_create_visit
    INDEX_LEVELS): every class is independent, each variable's term-sharing
    neighbours of lower index sit in earlier classes, and the sweeps end in
    the states, costs and accepted counts of the oracle's index-order sweeps
    (markov/walker.h:144-153); the lowest states follow the class order."""
    kind, gen = PROBLEMS[name]
    betas = np.linspace(0.1, 3.0, 24)
    op, qp, w, got, want = _run_pair(ctx, kind, gen(), R=37, B=B, betas=betas, visit=_qmc.VISIT_INDEX_LEVELS)
    order, co = qp.order(w.lanes())
    cls = np.zeros(op.n, dtype=int)
    for c in range(len(co) - 1):
        cls[order[co[c]:co[c + 1]]] = c
    off, ids, _ = op.edges()
    for j in range(len(off) - 1):
        m = sorted(ids[off[j]:off[j + 1]])
        for a, b in zip(m, m[1:]):
            assert cls[a] < cls[b]
    assert np.array_equal(got["lowest"], want["lowest"])
    assert np.array_equal(got["lowest_states"], want["lowest_states"])
    # the same sweeps in plain index order
    opts = po.make_opts(po.RNG_COUNTER, seed=7, arith=arith_of(qp))
    states = oracle_initial(op, opts, 37)
    costs = np.array([op.cost(s) for s in states])
    acc = op.run_sweeps(opts, states, costs, costs.copy(), states.copy(), np.arange(37), 0,
                        np.repeat(betas, 37))
    assert np.array_equal(got["states"], states)
    assert np.array_equal(got["cost"], costs)
    assert np.array_equal(got["accepted"], acc)


@pytest.mark.parametrize("lanes", [64, 128, 256, 512])
@pytest.mark.parametrize("B", [1, 8])
def test_sa_sweeps_lanes_and_slices(ctx, lanes, B):
    """Wider class rounds (several waves per replica group), ragged slices."""
    terms = wl.er_ising(1500, 7.0, seed=8)
    betas = np.linspace(0.1, 3.0, 10)
    op, qp, w, got, want = _run_pair(ctx, "ising", terms, R=9, B=B, betas=betas, lanes=lanes)
    assert w.lanes() == lanes
    for k in ("states", "cost", "lowest", "lowest_states", "accepted"):
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.parametrize("lanes", [64, 256])
@pytest.mark.parametrize("B", [2, 8])
def test_pubo_sliced_entries_lanes(ctx, lanes, B):
    """4-local PUBO through the sliced 16-byte entries at several lane counts."""
    terms = wl.random_pubo(700, 5000, seed=9)
    betas = np.linspace(0.01, 5.0, 10)
    op, qp, w, got, want = _run_pair(ctx, "pubo", terms, R=11, B=B, betas=betas, lanes=lanes)
    assert qp.info.pairwise == 0 and w.lanes() == lanes
    for k in ("states", "cost", "lowest", "lowest_states", "accepted"):
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.parametrize("name", ["lattice2d_pm1", "er_int", "pubo_int", "pubo_6local"])
def test_ordered_sweeps_bit_exact_int(ctx, name):
    kind, gen = PROBLEMS[name]
    betas = np.linspace(0.1, 3.0, 12)
    op, qp, w, got, want = _run_pair(ctx, kind, gen(), R=19, B=8, betas=betas, ordered=True)
    assert np.array_equal(got["states"], want["states"])
    assert np.array_equal(got["cost"], want["cost"])
    assert np.array_equal(got["lowest"], want["lowest"])
    assert np.array_equal(got["lowest_states"], want["lowest_states"])


def _bounded_degree(n, dmax, m, seed):
    """A ring plus random chords while both ends have fewer than dmax
    neighbours; integer couplings in +-[1, 3]."""
    rng = np.random.default_rng(seed)
    deg = np.zeros(n, dtype=int)
    edges = set()
    for i in range(n):
        edges.add((min(i, (i + 1) % n), max(i, (i + 1) % n)))
    deg += 2
    tries = 0
    while len(edges) < m and tries < 50 * m:
        tries += 1
        i, j = sorted(rng.choice(n, size=2, replace=False).tolist())
        if (i, j) in edges or deg[i] >= dmax or deg[j] >= dmax:
            continue
        edges.add((i, j))
        deg[i] += 1
        deg[j] += 1
    return [dict(c=int(rng.integers(1, 4)) * (1 if rng.random() < 0.5 else -1), ids=[i, j])
            for i, j in sorted(edges)]


ROW_CASES = {
    # widest slice -> entries read ahead per row: 2 -> 4 (padding), 4 -> 4,
    # 5 -> 6 (padding), 6 -> 6, > 8 -> 8 plus the rest per row
    "ring": lambda: _bounded_degree(300, 2, 300, seed=31),
    "deg5": lambda: _bounded_degree(400, 5, 900, seed=32),
    "lattice2d_pm1": lambda: wl.lattice2d(8, seed=3),
    "lattice2d_wide": lambda: [dict(t, c=t["c"] * 300) for t in wl.lattice2d(8, seed=3)],
    "lattice3d_int": lambda: [dict(t, c=int(np.sign(t["c"]) or 1)) for t in wl.lattice3d(6, seed=7)],
    "er_int": lambda: wl.er_ising(150, 6.0, seed=4),
}


@pytest.mark.parametrize("name", list(ROW_CASES))
@pytest.mark.parametrize("B", [1, 2])
def test_row_widths_bit_exact(ctx, name, B):
    """Kernels reading 4, 6 or 8 entries per pairwise row ahead (chosen from
    the widest slice; slices store at least 8 entries, zero-padded) against
    the oracle's sweeps, int8 and 4-byte-int entries both: identical states,
    costs, lowest costs and states, accepted counts."""
    betas = np.linspace(0.1, 3.0, 24)
    op, qp, w, got, want = _run_pair(ctx, "ising", ROW_CASES[name](), R=12, B=B, betas=betas)
    assert qp.info.coupling == _qmc.QMC_COUPLING_I32
    for k in ("states", "cost", "lowest", "lowest_states", "accepted"):
        assert np.array_equal(got[k], want[k]), k


def test_lowest_states_in_global_memory(ctx):
    """1024 work-groups of a 12000-variable graph: with the lowest states in
    LDS (three bytes per position) two work-groups fit a compute unit, with
    them in global memory three, so the launch keeps them in global memory;
    identical states, costs, lowest costs and states, accepted counts."""
    betas = np.linspace(0.1, 3.0, 4)
    op, qp, w, got, want = _run_pair(ctx, "ising", wl.er_ising(12000, 6.0, seed=41), R=1024, B=1, betas=betas,
                                     lanes=64)
    assert qp.info.coupling == _qmc.QMC_COUPLING_I32
    for k in ("states", "cost", "lowest", "lowest_states", "accepted"):
        assert np.array_equal(got[k], want[k]), k


def test_sa_sweeps_f32_six_entry_rows(ctx):
    """fp32 couplings on the 3D lattice (every slice 6 wide) with 2 replicas
    per work-group: the 6-entry kernel; identical states, costs within 1e-5."""
    kind, gen = PROBLEMS["lattice3d_f32"]
    betas = np.linspace(0.2, 3.0, 16)
    op, qp, w, got, want = _run_pair(ctx, kind, gen(), R=8, B=2, betas=betas)
    assert qp.info.coupling == _qmc.QMC_COUPLING_F32
    assert np.array_equal(got["states"], want["states"])
    assert np.array_equal(got["lowest_states"], want["lowest_states"])
    np.testing.assert_allclose(got["cost"], want["cost"], rtol=1e-5)
    np.testing.assert_allclose(got["lowest"], want["lowest"], rtol=1e-5)


def test_sa_sweeps_f32(ctx):
    kind, gen = PROBLEMS["lattice3d_f32"]
    betas = np.linspace(0.2, 3.0, 16)
    op, qp, w, got, want = _run_pair(ctx, kind, gen(), R=16, B=16, betas=betas)
    assert qp.info.coupling == _qmc.QMC_COUPLING_F32
    assert np.array_equal(got["states"], want["states"])
    e = np.array([op.cost(s) for s in got["lowest_states"]])
    np.testing.assert_allclose(w.energy(lowest=True), e, rtol=1e-12)
    np.testing.assert_allclose(got["lowest"], e, rtol=1e-5)
    np.testing.assert_allclose(got["cost"], want["cost"], rtol=1e-5)


@pytest.mark.parametrize("name,B,L,mode", [("er_int", 8, 8, 0), ("lattice2d_pm1", 16, 16, 0),
                                           ("er_int", 4, 8, 0), ("pubo_int", 8, 16, 0),
                                           ("er_int", 1, 8, 0), ("lattice2d_pm1", 2, 32, 0),
                                           ("er_int", 4, 8, 1), ("pubo_int", 8, 16, 1)])
def test_pt_matches_oracle(ctx, name, B, L, mode):
    kind, gen = PROBLEMS[name]
    terms = gen()
    op, qp = pair(ctx, kind, terms)
    temps = wl.geometric_temperatures(0.3, 3.0, L)
    betas = sorted([1.0 / t for t in temps], reverse=True)
    n_ladders = 3
    steps = 40
    w = _qmc.Walkers(qp, L * n_ladders, B, seed=21, replicas_per_seed=L)
    order, _ = qp.order(w.lanes())
    w.set_pt_mode(mode)
    w.init()
    w.pt(betas, 0, steps // 2)
    w.pt(betas, steps // 2, steps - steps // 2)  # continuation keeps the ladder
    got = w.read()
    labels = w.labels()
    for ld in range(n_ladders):
        opts = po.make_opts(po.RNG_COUNTER, seed=21 + ld, arith=arith_of(qp), order=order,
                            steps_per_tick=1)
        want = op.pt(opts, temps, False, steps)
        sl = slice(ld * L, (ld + 1) * L)
        assert np.array_equal(got["states"][sl], want["final_states"])
        assert np.array_equal(got["cost"][sl], want["final_costs"])
        assert np.array_equal(got["lowest"][sl], want["lowest"])
        assert np.array_equal(got["lowest_states"][sl], want["lowest_states"])
        assert np.allclose(got["beta"][sl], want["slot_beta"])


@pytest.mark.parametrize("side,n_ladders,steps", [(6, 3, 40), (16, 2, 24)])
def test_pt_f32_split_ladder_matches_oracle(ctx, side, n_ladders, steps):
    """The benchmark's kernel instantiation -- fp32 couplings, a 32-replica
    ladder split over 16 cooperative work-groups of 2 replicas, exchange
    inside the kernel through the cross-work-group energy hand-off -- step
    for step against the oracle's counter-mode parallel tempering
    (parallel_tempering.h:134-189) with the same fp32 field sums: identical
    states, lowest states and slot betas; costs and lowest costs within 1e-5
    relative (the device sums a class's accepted dE per wavefront).  The
    16^3 case is the benchmark problem at its default lane count."""
    terms = wl.lattice3d(side, seed=7)
    op, qp = pair(ctx, "ising", terms)
    assert qp.info.coupling == _qmc.QMC_COUPLING_F32
    L = 32
    temps = wl.geometric_temperatures(0.2, 2.0, L)
    betas = sorted([1.0 / t for t in temps], reverse=True)
    w = _qmc.Walkers(qp, L * n_ladders, 2, seed=31, replicas_per_seed=L)
    order, _ = qp.order(w.lanes())
    w.init()
    w.pt(betas, 0, steps)
    got = w.read()
    for ld in range(n_ladders):
        opts = po.make_opts(po.RNG_COUNTER, seed=31 + ld, arith=po.ARITH_F32, order=order, steps_per_tick=1)
        want = op.pt(opts, temps, False, steps)
        sl = slice(ld * L, (ld + 1) * L)
        assert np.array_equal(got["states"][sl], want["final_states"]), ld
        assert np.array_equal(got["lowest_states"][sl], want["lowest_states"]), ld
        assert np.allclose(got["beta"][sl], want["slot_beta"]), ld
        np.testing.assert_allclose(got["cost"][sl], want["final_costs"], rtol=1e-5)
        np.testing.assert_allclose(got["lowest"][sl], want["lowest"], rtol=1e-5)
        # every lowest state's fp64 energy is the lowest cost within 1e-5
        e = np.array([op.cost(s) for s in got["lowest_states"][sl]])
        np.testing.assert_allclose(got["lowest"][sl], e, rtol=1e-5)
    w.close()
    qp.close()


@pytest.mark.parametrize("B", [1, 4, 8])
def test_pt_observations_match_oracle(ctx, B):
    """Observed parallel tempering (qmc_walkers_pt_observe: graph-replayed
    steps whose exchange launch records each ladder position's cost,
    acceptances, direction label and pair swap) against the oracle's
    counter-mode run of the same ladders: per-position step averages of
    acc_rate, avg_cost, avg_dir and swap_rate as parallel_tempering.h:141-188
    observes them -- exact for integer couplings -- and the same final
    states as an unobserved run."""
    kind, gen = PROBLEMS["er_int"]
    op, qp = pair(ctx, kind, gen())
    L, n_ladders, steps = 8, 2, 30
    temps = wl.geometric_temperatures(0.3, 3.0, L)
    betas = sorted([1.0 / t for t in temps], reverse=True)
    w = _qmc.Walkers(qp, L * n_ladders, B, seed=21, replicas_per_seed=L)
    order, _ = qp.order(w.lanes())
    w.init()
    w.pt_observe(L, steps)
    w.pt(betas, 0, steps)
    got = w.read()
    log = w.pt_log()
    for ld in range(n_ladders):
        opts = po.make_opts(po.RNG_COUNTER, seed=21 + ld, arith=arith_of(qp), order=order, steps_per_tick=1)
        want = op.pt(opts, temps, False, steps)
        sl = slice(ld * L, (ld + 1) * L)
        assert np.array_equal(got["states"][sl], want["final_states"])
        rec = log[ld]
        assert np.array_equal(rec[:, :, 0].mean(axis=0), want["avg_cost"])
        assert np.allclose(rec[:, :, 1].mean(axis=0) / op.n, want["acc_rate"], rtol=1e-14, atol=0)
        assert np.array_equal(rec[:, :, 2].mean(axis=0), want["avg_dir"])
        sw = rec[:, :L - 1, 3]
        tried = sw >= 0
        assert np.array_equal(tried.sum(axis=0), [steps // 2] * (L - 1))
        assert np.allclose(np.where(tried, sw, 0).sum(axis=0) / tried.sum(axis=0), want["swap_rate"][:L - 1])


def test_batch_mixed_problems_same_as_alone(ctx):
    """Work-groups of different problems, sizes and kernel variants launched
    together, in chunks, give exactly the results of each run alone."""
    specs = [("ising", wl.er_ising(60, 6.0, seed=31), 16, 8, 3),
             ("ising", wl.er_ising(300, 6.0, seed=32), 24, 8, 4),
             ("pubo", wl.random_pubo(100, 400, seed=33), 8, 8, 5),
             ("ising", wl.lattice2d(8, seed=34), 32, 32, 6)]
    sweeps = 30
    alone, mixed, keep = [], [], []
    for kind, terms, R, B, seed in specs:
        op, qp = pair(ctx, kind, terms)
        keep.append(qp)
        betas = np.linspace(0.2, 2.5, sweeps) * (1 + 0.1 * seed)
        w = _qmc.Walkers(qp, R, B, seed=seed)
        w.init()
        w.sweep(0, schedule=betas)
        alone.append(w.read())
        w2 = _qmc.Walkers(qp, R, B, seed=seed)
        w2.init()
        w2.plan(0, sweeps, betas)
        mixed.append(w2)
    batch = _qmc.Batch(ctx)
    for w2 in mixed:
        batch.add(w2)
    for _ in range(4):
        batch.run(8)
        batch.wait()
    assert all(w2.remaining() == 0 for w2 in mixed)
    for a, w2 in zip(alone, mixed):
        b = w2.read()
        for k in ("states", "cost", "lowest", "lowest_states", "accepted"):
            assert np.array_equal(a[k], b[k]), k
    # one batched read-back of every member: energies recomputed from the
    # lowest states (integer couplings: equal to the tracked lowest costs)
    for a, w2, r in zip(alone, mixed, batch.read()):
        assert np.array_equal(r["lowest"], a["lowest"])
        assert np.array_equal(r["lowest"], w2.energy(lowest=True))
        assert np.array_equal(r["accepted"], a["accepted"])
        assert np.array_equal(r["lowest_states"], a["lowest_states"])


def test_batch_groups_on_several_streams(ctx):
    """Members with different work-group sizes (lane counts) and kernel
    variants run as separate launches spread over several streams; each
    result equals its run alone."""
    specs = [("ising", wl.er_ising(900, 6.0, seed=41), 8, 8, 512),
             ("ising", wl.er_ising(400, 6.0, seed=42), 8, 8, 128),
             ("ising", wl.er_ising(200, 6.0, seed=43), 8, 1, 64),
             ("pubo", wl.random_pubo(150, 600, seed=44), 4, 4, 64),
             ("ising", wl.lattice3d(8, seed=45), 16, 2, 256)]
    sweeps = 40
    alone, mixed, keep = [], [], []
    for i, (kind, terms, R, B, lanes) in enumerate(specs):
        op, qp = pair(ctx, kind, terms)
        keep.append(qp)
        betas = np.linspace(0.1, 3.0, sweeps)
        w = _qmc.Walkers(qp, R, B, seed=100 + i)
        w.set_lanes(lanes)
        w.init()
        w.sweep(0, schedule=betas)
        alone.append(w.read())
        w2 = _qmc.Walkers(qp, R, B, seed=100 + i)
        w2.set_lanes(lanes)
        w2.init()
        w2.plan(0, sweeps, betas)
        mixed.append(w2)
    batch = _qmc.Batch(ctx)
    for w2 in mixed:
        batch.add(w2)
    batch.run(1 << 20)
    batch.wait()
    for a, w2, r in zip(alone, mixed, batch.read()):
        b = w2.read()
        for k in ("states", "cost", "lowest", "lowest_states", "accepted"):
            assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(r["lowest_states"], a["lowest_states"])
    # the same plans as graph-replayed launches of 7 sweeps (replayed twice:
    # re-planned walkers reuse the recorded graph)
    for rep in range(2):
        acc0 = []
        for i, w2 in enumerate(mixed):
            acc0.append(w2.read()["accepted"])
            w2.init()
            w2.plan(0, sweeps, np.linspace(0.1, 3.0, sweeps))
        batch.run_chunked(7)
        batch.wait()
        for a, w2, a0 in zip(alone, mixed, acc0):
            b = w2.read()
            assert w2.remaining() == 0
            for k in ("states", "cost", "lowest", "lowest_states"):
                assert np.array_equal(a[k], b[k]), (rep, k)
            # accepted counters accumulate over runs
            assert np.array_equal(a["accepted"], b["accepted"] - a0), rep


@pytest.mark.parametrize("kind,terms,n_ladders,steps,groups",
                         [("ising", "lattice3d", 6, 300, (32, 16, 8, 4, 2, 1)),
                          ("ising", "er", 6, 300, (32, 8, 1)),
                          ("ising", "lattice16", 16, 1500, (32, 2, 1))])
def test_pt_independent_of_replica_grouping(ctx, kind, terms, n_ladders, steps, groups):
    """Cooperative ladders: the same run for every split of a 32-replica
    ladder over work-groups (the RNG contract does not see the grouping), so
    the cross-work-group energy hand-off cannot have lost an update.  The
    last case is the full-size benchmark problem with 512 work-groups."""
    tr = {"lattice3d": lambda: wl.lattice3d(8, seed=12), "er": lambda: wl.er_ising(700, 6.0, seed=13),
          "lattice16": lambda: wl.lattice3d(16, seed=14)}[terms]()
    op, qp = pair(ctx, kind, tr)
    temps = wl.geometric_temperatures(0.2, 2.0, 32)
    betas = sorted([1.0 / t for t in temps], reverse=True)
    ref = None
    for B in groups:
        w = _qmc.Walkers(qp, 32 * n_ladders, B, seed=5, replicas_per_seed=32)
        w.set_lanes(128)
        w.init()
        w.pt(betas, 0, steps)
        got = w.read()
        got["labels"] = w.labels()
        if ref is None:
            ref = got
            continue
        for k in ("states", "cost", "lowest", "labels", "beta"):
            assert np.array_equal(got[k], ref[k]), (B, k)


def test_handles_destroyed_in_any_order():
    """Handles are reference counted: closing the context or the problem
    first leaves walkers and batches usable until they are destroyed."""
    c = _qmc.Context(0)
    op, qp = pair(c, "ising", wl.er_ising(120, 6.0, seed=51))
    w = _qmc.Walkers(qp, 8, 8, seed=3)
    w.init()
    betas = np.linspace(0.1, 3.0, 10)
    w.plan(0, 10, betas)
    b = _qmc.Batch(c)
    b.add(w)
    c.close()
    qp.close()
    b.run(1 << 20)
    b.wait()
    got = w.read()
    assert got["lowest"].shape == (8,)
    w.close()
    assert b.read()[0]["lowest"].shape == (8,)
    b.close()


def test_lds_bandwidth_probe(ctx):
    """The LDS streaming probe reports a plausible aggregate bandwidth
    (MI355X: 256 CUs x 128 B/clk x 2.4 GHz = 78.6 TB/s nominal)."""
    bw = ctx.lds_bandwidth()
    assert 1e13 < bw < 2e14, bw


def test_row_entry_widths_and_wide_rows(ctx):
    """Pairwise rows hold 4-byte entries (24-bit position, int8 coefficient)
    for integer coefficients in [-128, 127] and 8-byte entries otherwise (fp32,
    or an integer beyond int8); both forms are bit-exact against the oracle
    (the 4-byte form through every integer pairwise test; here the 8-byte form
    with an integer coefficient of 1000, and an fp32 ER model: identical
    states, costs within 1e-5)."""
    kind, gen = PROBLEMS["lattice3d_f32"]
    _, qp = pair(ctx, kind, gen())
    assert qp.info.entry_bytes == 8
    _, qp = pair(ctx, "ising", wl.er_ising(150, 6.0, seed=4))
    assert qp.info.entry_bytes == 4
    big = wl.er_ising(150, 6.0, seed=4)
    big[-1] = dict(big[-1], c=1000)
    op, qp, w, got, want = _run_pair(ctx, "ising", big, R=16, B=8, betas=np.linspace(0.1, 3.0, 12))
    assert qp.info.entry_bytes == 8 and qp.info.coupling == _qmc.QMC_COUPLING_I32
    for k in ("states", "cost", "lowest", "lowest_states"):
        assert np.array_equal(got[k], want[k]), k
    _, qp = pair(ctx, "pubo", wl.random_pubo(120, 600, seed=5))
    assert qp.info.entry_bytes == 16
    terms = wl.er_ising_f32(5000, 8.0, seed=17)
    op, qp, w, got, want = _run_pair(ctx, "ising", terms, R=16, B=16, betas=np.linspace(0.2, 3.0, 12))
    assert qp.info.entry_bytes == 8 and qp.info.coupling == _qmc.QMC_COUPLING_F32
    assert np.array_equal(got["states"], want["states"])
    np.testing.assert_allclose(got["cost"], want["cost"], rtol=1e-5)
    np.testing.assert_allclose(got["lowest"], want["lowest"], rtol=1e-5)
