"""The serving loop on the GPU: requests of different kinds and sizes,
submitted together or staggered, complete with exactly the result each
gets when solved alone (integer couplings bit-exact; fp32 within 1e-5
relative), admission respects the HBM budget, and oversized requests fail
with MEMORY_LIMITED."""
import time

import numpy as np
import pytest

from qiotoolkit_amd import workloads as wl

pytestmark = pytest.mark.gpu

SA = "simulatedannealing.qiotoolkit"
PA = "populationannealing.cpu"


def pubo_arrays(n_vars, n_terms, seed):
    terms = wl.random_pubo(n_vars, n_terms, seed=seed)
    off = np.cumsum([0] + [len(t["ids"]) for t in terms]).astype(np.int64)
    ids = np.array([i for t in terms for i in t["ids"]], dtype=np.int32)
    c = np.array([t["c"] for t in terms], dtype=np.float64)
    return off, ids, c


def f32_arrays(n, seed):
    terms = wl.er_ising_f32(n, 8.0, seed=seed)
    off = np.cumsum([0] + [len(t["ids"]) for t in terms]).astype(np.int64)
    ids = np.array([i for t in terms for i in t["ids"]], dtype=np.int32)
    c = np.array([t["c"] for t in terms], dtype=np.float64)
    return off, ids, c


def requests():
    out = []
    for k, n in enumerate([300, 1200, 90]):
        out.append((SA, {"beta_start": 0.1, "beta_stop": 3.0, "restarts": 16, "sweeps": 120, "seed": 5 + k},
                    "ising", wl.er_ising_arrays(n, 6.0, 3, seed=80 + k)))
    out.append((PA, {"population": 64, "beta_start": 0, "beta_stop": 3, "sweeps": 20,
                     "sweeps_per_replica": 4, "seed": 21}, "ising", wl.er_ising_arrays(500, 8.0, 3, seed=90)))
    out.append((PA, {"population": 48, "beta_start": 0, "beta_stop": 3, "sweeps": 15,
                     "sweeps_per_replica": 3, "seed": 22}, "pubo", pubo_arrays(200, 1000, seed=91)))
    out.append((SA, {"beta_start": 0.1, "beta_stop": 3.0, "restarts": 8, "sweeps": 80, "seed": 23},
                "ising", f32_arrays(400, seed=92)))
    return out


def alone_results(reqs):
    from qiotoolkit_amd.solver import Engine
    eng = Engine(0)
    res = []
    for target, params, kind, arrays in reqs:
        m = eng.model_from_arrays(kind, *arrays)
        res.append(m.solve(target, {"params": params}))
        m.close()
    eng.close()
    return res


def check_same(a, b, fp32=False):
    ca, cb = a["solutions"]["cost"], b["solutions"]["cost"]
    if fp32:
        assert abs(ca - cb) <= 1e-5 * max(1.0, abs(ca))
    else:
        assert ca == cb
    assert a["solutions"]["configuration"] == b["solutions"]["configuration"]


@pytest.mark.parametrize("stagger_ms", [0, 30])
def test_results_same_alone_and_in_the_mix(stagger_ms):
    from qiotoolkit_amd.solver import Server
    reqs = requests()
    alone = alone_results(reqs)
    srv = Server(0, max_active=8, sweeps_per_round=7)
    ids = []
    for target, params, kind, arrays in reqs:
        ids.append(srv.submit(target, {"params": params}, kind, *arrays))
        if stagger_ms:
            time.sleep(stagger_ms * 1e-3)
    got = []
    for rid in ids:
        rc, doc = srv.wait(rid, 120000)
        assert rc == 0, doc
        got.append(doc)
    st = srv.stats()
    srv.close()
    for k, (a, b) in enumerate(zip(alone, got)):
        check_same(a, b, fp32=(k == len(reqs) - 1))
        sv = b["benchmark"]["serving"]
        assert sv["finished_ms"] >= sv["admitted_ms"] >= sv["arrival_ms"]
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["latency_p99_ms"] >= st["latency_p50_ms"] > 0


def test_admission_by_device_budget():
    """A budget below the three requests' footprints together: every
    request still completes with its alone result, never more than the
    budget is reserved, and a request larger than the whole budget fails
    with MEMORY_LIMITED."""
    from qiotoolkit_amd.solver import Server
    reqs = requests()[:3]
    alone = alone_results(reqs)
    srv = Server(0, max_active=8)
    fps = []
    for t, p, k, a in reqs:
        rc, doc = srv.wait(srv.submit(t, {"params": p}, k, *a), 120000)
        fps.append(doc["benchmark"]["serving"]["device_bytes_reserved"])
    srv.close()
    budget = max(fps) + min(fps)
    assert budget < sum(fps)
    srv = Server(0, max_active=8, sweeps_per_round=10, hbm_budget_bytes=budget)
    ids = [srv.submit(t, {"params": p}, k, *a) for t, p, k, a in reqs]
    big = srv.submit(SA, {"params": {"restarts": 256, "sweeps": 10, "beta_start": 0.1, "beta_stop": 1.0}},
                     "ising", *wl.er_ising_arrays(20000, 6.0, 3, seed=7))
    got = [srv.wait(i, 120000) for i in ids]
    rc_big, doc_big = srv.wait(big, 120000)
    st = srv.stats()
    srv.close()
    for a, (rc, b) in zip(alone, got):
        assert rc == 0
        check_same(a, b)
    assert rc_big == 1 and doc_big["error"]["code"] == 1
    assert 0 < st["peak_reserved_bytes"] <= budget
    assert st["peak_active"] <= 2


def test_limited_requests_run_as_rounds():
    """Requests with a time or cost limit are advanced round by round like
    the others: a time-limited SA request of a million sweeps completes at its
    limit with the reference's exit reason while the unlimited requests
    submitted with it finish long before, each with its alone result; a
    cost-limited PA request stops at its first check."""
    from qiotoolkit_amd.solver import Server
    reqs = requests()[:3]
    alone = alone_results(reqs)
    srv = Server(0, max_active=8, sweeps_per_round=10)
    long_id = srv.submit(SA, {"params": {"beta_start": 0.1, "beta_stop": 3.0, "restarts": 8, "sweeps": 1000000,
                                         "timeout": 2.0, "seed": 9}},
                         "ising", *wl.er_ising_arrays(400, 6.0, 3, seed=77))
    cost_id = srv.submit(PA, {"params": {"population": 32, "beta_start": 0, "beta_stop": 3, "sweeps": 2000,
                                         "cost_limit": 1e9, "seed": 4}},
                         "ising", *wl.er_ising_arrays(300, 6.0, 3, seed=78))
    ids = [srv.submit(t, {"params": p}, k, *a) for t, p, k, a in reqs]
    got = [srv.wait(i, 120000) for i in ids]
    rc_long, doc_long = srv.wait(long_id, 120000)
    rc_cost, doc_cost = srv.wait(cost_id, 120000)
    srv.close()
    for a, (rc, b) in zip(alone, got):
        assert rc == 0
        check_same(a, b)
        # not held up behind the long request
        assert b["benchmark"]["serving"]["finished_ms"] < doc_long["benchmark"]["serving"]["finished_ms"] - 500
    assert rc_long == 0
    sol = {e["name"]: e["value"] for e in doc_long["benchmark"]["extensions"]}["solver"]
    assert sol["exit_reason"].startswith("Stop due to time_limit.steps : ")
    assert 0 < sol["last_step"] < 1000000
    lat = doc_long["benchmark"]["serving"]["finished_ms"] - doc_long["benchmark"]["serving"]["admitted_ms"]
    assert 2000 <= lat < 20000
    assert rc_cost == 0
    sol = {e["name"]: e["value"] for e in doc_cost["benchmark"]["extensions"]}["solver"]
    assert sol["exit_reason"].startswith("Stop due to cost limit:")
    assert sol["last_step"] < 2000
