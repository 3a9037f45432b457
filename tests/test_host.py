"""Host-side logic of the product (no GPU): libraries load and export every
declared symbol, and the C++ host layer agrees with the oracle on
normalisation, schedules and the resolved solver plans."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(q(?:mc|iotk)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,libname", [("qmc.h", "libqmc.so"), ("qiotk.h", "libqiotk.so")])
def test_library_exports_every_declared_symbol(header, libname):
    from qiotoolkit_amd import _qmc, solver  # noqa: F401  (loads libqmc first)
    _qmc.lib()
    lib = ctypes.CDLL(os.path.join(ROOT, "qiotoolkit_amd", libname))
    names = declared(header)
    assert len(names) > 5
    for n in names:
        assert hasattr(lib, n), n
    listed = _qmc.EXPORTS if header == "qmc.h" else solver.EXPORTS
    assert sorted(listed) == names


def test_no_device_fails_loudly():
    from qiotoolkit_amd import _qmc
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(_qmc.QmcError):
        _qmc.Context(0)
    from qiotoolkit_amd import solver as S
    with pytest.raises(S.SolverError):
        S.Engine(0)
    with pytest.raises(S.SolverError):
        S.Server(0, max_active=4)


def load(name):
    return json.load(open(os.path.join(GOLD, name)))


@pytest.mark.parametrize("fname", ["ising1.json", "cpupubojobstest.json", "isingemptymixed.json",
                                   "puboemptymixed.json", "isingempty.json",
                                   "e2e/combined_4_local_N_16_qualitytest/combined_4_local_N_16_inst_1.json"])
def test_normalisation_matches_oracle(fname):
    from qiotoolkit_amd import solver as S
    doc = load(fname)
    cf = doc["cost_function"]
    m = S.parse_model(doc)
    op = po.Problem(cf["type"], cf["terms"])
    assert m["n_vars"] == op.n
    assert m["const_cost"] == op.const_cost
    if op.n:
        assert m["names"] == op.names().tolist()
        off, ids, c = op.edges()
        assert m["term_offsets"] == off.tolist()
        assert m["term_vars"] == ids.tolist()
        assert m["coeffs"] == c.tolist()


def test_normalisation_errors():
    from qiotoolkit_amd import solver as S
    with pytest.raises(S.SolverError) as e:
        S.parse_model({"cost_function": {"type": "ising", "terms": [{"c": 1, "ids": [3, 3]}]}})
    assert e.value.code == 101 and "Duplicated" in e.value.message  # utils::Error::DuplicatedVariable
    m = S.parse_model({"cost_function": {"type": "pubo", "terms": [{"c": 1, "ids": [3, 3, 4]}]}})
    assert m["term_vars"] == [0, 1]
    with pytest.raises(S.SolverError, match="greater than 0"):
        S.parse_model({"cost_function": {"type": "ising", "terms": []}})
    with pytest.raises(S.SolverError, match="version"):
        S.parse_model({"cost_function": {"type": "ising", "version": "-7", "terms": [{"c": 1, "ids": [0]}]}})
    with pytest.raises(S.SolverError, match="not found"):
        S.parse_model({"cost_function": {"type": "ising", "terms": [{"c": 1, "ids": [0, 1]}],
                                         "initial_configuration": {"2": 1}}})


def test_schedules_known_values():
    from qiotoolkit_amd import solver as S
    y, _ = S.schedule_eval({"type": "linear", "initial": 3, "final": 5}, [-1, 0, 0.5, 1, 2])
    assert y.tolist() == [3, 3, 4, 5, 5]
    y, _ = S.schedule_eval({"type": "geometric", "initial": 3, "final": 6}, [0, 0.5, 1])
    assert y[0] == 3 and y[2] == 6 and abs(y[1] - 3 * 2 ** 0.5) < 1e-15
    _, meta = S.schedule_eval({"type": "geometric", "initial": 10, "final": 1, "count": 100}, [0])
    assert abs(meta["values"][70] - 1.9630406500402637) < 1e-14
    y, meta = S.schedule_eval({"type": "linear", "initial": 3, "final": 5, "start": 2, "stop": 4,
                               "repeat": True}, [1.5, 2.5, 6.5])
    assert y.tolist() == [4.5, 3.5, 3.5]  # schedule/test/linear_test.cc:83-92
    _, meta = S.schedule_eval([0.1, 0.5, 1, 2, 4], [0])
    assert meta["values"] == [0.1, 0.5, 1, 2, 4]


@pytest.mark.parametrize("b0,b1,sweeps", [(0.1, 3.0, 1000), (0.01, 5.0, 2000), (0.5, 1.0, 3)])
def test_sa_plan_matches_oracle_schedule(b0, b1, sweeps):
    from qiotoolkit_amd import solver as S
    d = S.describe("simulatedannealing.qiotoolkit",
                   {"params": {"beta_start": b0, "beta_stop": b1, "sweeps": sweeps, "restarts": 7}})
    steps = max(sweeps, 10)
    assert d["steps"] == steps and d["restarts"] == 7
    stop = max(1, sweeps - 1)
    want = [po.lib().orc_schedule_value(0, b0, b1, 0, stop, s) for s in range(steps)]
    assert d["betas"] == want


def test_sa_plan_errors():
    from qiotoolkit_amd import solver as S
    with pytest.raises(S.SolverError, match="greater than 1"):
        S.describe("simulatedannealing.qiotoolkit", {"params": {"sweeps": 1}})
    with pytest.raises(S.SolverError, match="beta_stop"):
        S.describe("simulatedannealing.qiotoolkit", {"params": {"beta_start": 2, "beta_stop": 1}})
    with pytest.raises(S.SolverError, match="Unknown solver"):
        S.describe("tabu.qiotoolkit", {"params": {}})


def test_pt_plan_ladder():
    from qiotoolkit_amd import solver as S
    d = S.describe("paralleltempering.qiotoolkit",
                   {"params": {"temperatures": {"type": "geometric", "initial": 0.2, "final": 2.0,
                                                "count": 32}, "sweeps": 5000}})
    t = np.zeros(32)
    import ctypes as C
    po.lib().orc_schedule_discretized(1, 0.2, 2.0, 32, t.ctypes.data_as(C.POINTER(C.c_double)))
    assert d["betas"] == [1.0 / x for x in sorted(t)]
    d = S.describe("paralleltempering.qiotoolkit", {"params": {"sweeps": 100}})
    assert d["betas"] == [4, 2, 1, 0.5, 0.1]  # default all_betas, descending


def test_error_codes_are_the_references():
    """Codes of utils::Error (utils/error_handling.h:25-52) for the
    reference's own messages: DuplicatedVariable 101 (graph/graph.h:130),
    MissingInput 102 (model/base_model.cc:8-14), ValueError 107
    (model/ising.h:86-87), ParsingError 105 for malformed text."""
    import ctypes as C
    from qiotoolkit_amd import solver as S
    cases = [
        ({"cost_function": {"type": "ising", "terms": [{"c": 1, "ids": [3, 3]}]}}, 101),
        ({"not_a_cost_function": {}}, 102),
        ({"cost_function": {"type": "ising", "version": "9.9", "terms": [{"c": 1, "ids": [0, 1]}]}}, 107),
    ]
    for doc, code in cases:
        with pytest.raises(S.SolverError) as e:
            S.parse_model(doc)
        assert e.value.code == code, (doc, e.value.code, e.value.message)
    out = C.c_void_p()
    rc = S.lib().qiotk_parse_model(b'{"cost_function": {"type": "ising", "terms": [', C.byref(out))
    assert rc == 105
    S.lib().qiotk_free(out)
    with pytest.raises(S.SolverError) as e:
        S.describe("simulatedannealing.qiotoolkit", {"params": {"sweeps": 1}})
    assert e.value.code == 107


def test_pa_configure_strategies_and_errors():
    """PopulationAnnealing::configure (population_annealing.h:622-751):
    every resampling_strategy is accepted; an unknown one is InitialConfig
    104; the constants and the population must be positive (ValueError 107,
    the matcher's wording, population_annealing_test.cc:213-230)."""
    from qiotoolkit_amd import solver as S
    for strat in ("linear_schedule", "geometric_schedule", "friction_tensor", "energy_variance",
                  "constant_culling"):
        d = S.describe("populationannealing.cpu",
                       {"params": {"resampling_strategy": strat, "sweeps": 50, "constant_population": True}})
        assert d["steps"] == 50
    cases = [
        ({"resampling_strategy": "magic"}, 104, "resampling strategy` must be one of"),
        ({"resampling_strategy": "constant_culling", "culling_fraction": 0}, 107,
         "parameter `culling_fraction`: must be greater than 0, found 0"),
        ({"resampling_strategy": "friction_tensor", "friction_tensor_constant": -1}, 107,
         "friction_tensor_constant"),
        ({"resampling_strategy": "energy_variance", "initial_culling_fraction": 0}, 107,
         "initial_culling_fraction"),
        ({"population": 0, "resampling_strategy": "constant_culling", "culling_fraction": 0.000001}, 107,
         "parameter `population`: must be greater than 0, found 0"),
        ({"beta_start": 4, "beta_stop": 2}, 107, "beta"),
        ({"beta_start": 0, "beta_stop": 2, "resampling_strategy": "geometric_schedule"}, 107, "positive"),
    ]
    for params, code, text in cases:
        with pytest.raises(S.SolverError) as e:
            S.describe("populationannealing.cpu", {"params": dict(params, sweeps=50)})
        assert e.value.code == code and text in e.value.message, (params, e.value.code, e.value.message)


# ------------------------------------------------- protobuf problem folders

PROTO = os.path.join(os.path.dirname(__file__), "golden", "proto")


def _ring(n, coeffs, kind):
    return {"cost_function": {"type": kind, "version": "1.0",
                              "terms": [{"c": coeffs[i], "ids": [i, (i + 1) % n]} for i in range(n)]}}


def _model(desc):
    desc.pop("input_reader", None)
    return desc


def test_proto_folders_of_the_reference():
    """The reference's own .pb fixtures (data/*_pb) read as the equivalent
    JSON documents its tests compare them with (model/test/test_utils.h:10-52,
    utils/test/stream_proto_test.cc:89-103)."""
    from qiotoolkit_amd import solver as S
    m = S.parse_model_file(os.path.join(PROTO, "input_problem_pb"))  # 4 messages
    assert m["kind"] == "ising" and len(m["coeffs"]) == 10
    assert m["coeffs"][0] == 2.5 and m["term_vars"][:2] == [0, 1]
    assert m["coeffs"] == [2.5, 2.5, 3.5, 2.5, 3.5, 4.5, 2.5, 3.5, 4.5, 5.5]
    ring = _model(S.parse_model_file(os.path.join(PROTO, "models_input_problem_pb")))
    assert ring == _model(S.parse_model(_ring(10, [1.0] * 10, "pubo")))
    c2 = [-9.5 if i == 9 else (-(i + 1.0) if i % 2 == 0 else i + 1.0) for i in range(10)]
    assert _model(S.parse_model_file(os.path.join(PROTO, "models_input_problem2_pb"))) == _model(S.parse_model(_ring(10, c2, "ising")))
    assert _model(S.parse_model_file(os.path.join(PROTO, "graph_input_problem_pb"))) == _model(S.parse_model(_ring(4, [1.0] * 4, "ising")))
    with pytest.raises(S.SolverError):  # soft-spin problems are not on this engine
        S.parse_model_file(os.path.join(PROTO, "ex_input_problem_pb"))


def _pb_folder(tmp_path, name, payloads):
    d = tmp_path / name
    d.mkdir()
    for k, b in enumerate(payloads):
        (d / ("%s_%d.pb" % (name, k))).write_bytes(b)
    return str(d)


def _len_field(field, payload):
    assert len(payload) < 128
    return bytes([field << 3 | 2, len(payload)]) + payload


def test_proto_folder_errors(tmp_path):
    """utils/test/stream_proto_test.cc:33-86, 106-116: the reference's
    messages and its utils::Error codes."""
    from qiotoolkit_amd import solver as S
    missing = str(tmp_path / "nowhere.json")  # not a folder: read as a JSON file
    with pytest.raises(S.SolverError) as e:
        S.parse_model_file(missing)
    assert e.value.code == 301 and e.value.message == "Could not open file: " + missing
    empty = tmp_path / "err_input_problem_pb"
    empty.mkdir()
    with pytest.raises(S.SolverError) as e:
        S.parse_model_file(str(empty))
    assert e.value.code == 301
    assert e.value.message == "Could not open file: " + str(empty) + "/err_input_problem_pb_0.pb"
    ising_foo = _len_field(1, bytes([0x08, 0x01]) + _len_field(2, b"foo"))
    with pytest.raises(S.SolverError) as e:
        S.parse_model_file(_pb_folder(tmp_path, "a_pb", [ising_foo]))
    assert e.value.code == 107 and e.value.message == "Expected version to be 1.0 or 1.1. Provided: foo"
    with pytest.raises(S.SolverError) as e:
        S.parse_model_file(_pb_folder(tmp_path, "b_pb", [b""]))
    assert e.value.code == 102 and e.value.message == "Invalid problem message. No cost function found"
    ising_no_terms = _len_field(1, bytes([0x08, 0x01]) + _len_field(2, b"1.0"))
    with pytest.raises(S.SolverError) as e:
        S.parse_model_file(_pb_folder(tmp_path, "c_pb", [ising_no_terms]))
    assert e.value.code == 107 and e.value.message == "Problem terms cannot be 0. Please add problem terms"
    # unpacked ids and a term split over two messages read the same as packed ones
    import struct
    term = lambda c, ids: _len_field(3, bytes([0x09]) + struct.pack("<d", c) + b"".join(bytes([0x10, i]) for i in ids))
    first = _len_field(1, bytes([0x08, 0x01]) + _len_field(2, b"1.0") + term(-1.5, [0, 2]))
    second = _len_field(1, term(2.0, [2, 5]))
    m = _model(S.parse_model_file(_pb_folder(tmp_path, "d_pb", [first, second])))
    assert m == _model(S.parse_model({"cost_function": {"type": "ising", "version": "1.0", "terms": [
        {"c": -1.5, "ids": [0, 2]}, {"c": 2.0, "ids": [2, 5]}]}}))


def test_json_input_file(tmp_path):
    from qiotoolkit_amd import solver as S
    doc = _ring(6, [1.0, -2.0, 3.0, -4.0, 5.0, -6.0], "ising")
    p = tmp_path / "input.json"
    p.write_text(json.dumps(doc))
    assert _model(S.parse_model_file(str(p))) == _model(S.parse_model(doc))


def _step_events(steps, per_step, restart_at=()):
    ev = []
    for s in range(steps):
        ev += [{"label": "step", "value": s}, {"observe": "step", "value": s}]
        ev += per_step(s)
        ev.append({"label": "step", "value": None})
        if s in restart_at:
            ev.append({"restart": s})
    return ev


def test_observer_windows_groups_and_kinds():
    """observe/observer.cc: average_over windows start at the last restart
    (v - (v - restart) % w), group_by labels nest inside them in label
    order, Constant keeps the last value of positive weight, Average the
    weighted mean; groups render as arrays in label-value order."""
    from qiotoolkit_amd import solver as S

    def per_step(s):
        ev = []
        for r in range(2):
            ev += [{"label": "replica", "value": r}, {"observe": "T", "value": 0.1 * (r + 1)},
                   {"observe": "avg_cost", "value": s + r}, {"label": "replica", "value": None}]
        return ev
    cfg = {"observer": {"average_over": {"step": 2}, "group_by": ["replica"], "constant": ["T"],
                        "average": ["avg_cost", "step"]}}
    out = S.observe(cfg, _step_events(7, per_step, restart_at=(3,)))
    # windows: {0,1} {2,3} | restart at 3 | {4} {5,6}
    assert [g["step"] for g in out["step"]] == [0.5, 2.5, 4, 5.5]
    assert [[r["avg_cost"] for r in g["replica"]] for g in out["step"]] == [[0.5, 1.5], [2.5, 3.5], [4, 5],
                                                                            [5.5, 6.5]]
    assert all([r["T"] for r in g["replica"]] == [0.1, 0.2] for g in out["step"])
    # weights, zero-weight constants, and a protocol list
    ev = [{"observe": "a", "value": 1.0, "weight": 3}, {"observe": "a", "value": 5.0},
          {"observe": "c", "value": 2.0}, {"observe": "c", "value": 9.0, "weight": 0},
          {"observe": "u", "value": 4.0}]
    out = S.observe({"observer": [{"average": ["a"]}, {"constant": ["c"]}]}, ev)
    assert out == {"a": 2.0, "c": 2.0}


def test_observer_pt_render_shape_and_errors():
    """The PT observation sequence of parallel_tempering.h:141-186 renders as
    the reference's toy test prints it (parallel_tempering_test.cc:153-190):
    one entry per replica with T, acc_rate, avg_cost, avg_dir, and
    swap_rate for every replica but the last."""
    from qiotoolkit_amd import solver as S
    L = 5

    def per_step(s):
        ev = []
        for i in range(L):
            ev += [{"label": "replica", "value": i}, {"observe": "T", "value": 0.1 * (i + 1)},
                   {"observe": "acc_rate", "value": 0.5}, {"observe": "avg_cost", "value": 1.0},
                   {"observe": "avg_dir", "value": i % 2}, {"label": "replica", "value": None}]
        for i in range(s & 1, L - 1, 2):
            ev += [{"label": "replica", "value": i}, {"observe": "swap_rate", "value": (i + s) % 2},
                   {"label": "replica", "value": None}]
        return ev
    cfg = {"observer": {"group_by": ["replica"], "constant": ["T"],
                        "average": ["avg_cost", "acc_rate", "avg_dir", "swap_rate"]}}
    out = S.observe(cfg, _step_events(200, per_step))
    reps = out["replica"]
    assert len(reps) == L and [r["T"] for r in reps] == [0.1, 0.2, 0.30000000000000004, 0.4, 0.5]
    assert all(set(r) == {"T", "acc_rate", "avg_cost", "avg_dir", "swap_rate"} for r in reps[:-1])
    assert set(reps[-1]) == {"T", "acc_rate", "avg_cost", "avg_dir"}
    with pytest.raises(S.SolverError) as e:
        S.observe({"observer": [3]}, [])
    assert e.value.code == 103
    S.describe("paralleltempering.qiotoolkit", {"params": {"sweeps": 10}, "observer": {"average": ["acc_rate"]}})
    S.describe("paralleltempering.qiotoolkit", {"params": {"sweeps": 10}, "observer": {"constant": ["beta"]}})


def _ring_doc(kind, coeffs, extra=()):
    terms = [{"c": c, "ids": [i, (i + 1) % len(coeffs)]} for i, c in enumerate(coeffs)] + list(extra)
    return {"cost_function": {"type": kind, "version": "1.1", "terms": terms}}


@pytest.mark.parametrize("doc,mn,mx", [
    # ring-10 Ising of unit couplings: every variable's two |c| cancel, so the
    # default 0.1 doubled; max 2 x 2 (ising_test.cc:162-168)
    (_ring_doc("ising", [1] * 10), 0.2, 4.0),
    # alternating ring: least difference 0.5 at variable 9, doubled (ising_test.cc:407-432)
    (_ring_doc("ising", [1, -2, 3, -4, 5, -6, 7, -8, 9, -9.5]), 1.0, None),
    # ring-10 PUBO: no negative coefficients, so the least positive one (pubo_test.cc:163-169)
    (_ring_doc("pubo", [1] * 10), 1.0, 2.0),
    # pubo_test.cc:364-379: +2 and -2 on one variable cancel; max |-2| + |2|
    ({"cost_function": {"type": "pubo", "version": "1.1", "terms": [
        {"c": 1, "ids": []}, {"c": -2, "ids": [1, 2]}, {"c": 2, "ids": [1, 2]}]}}, 0.1, 4.0),
    # pubo_test.cc:381-410
    ({"cost_function": {"type": "pubo", "version": "1.1", "terms": [
        {"c": 1, "ids": []}, {"c": -2, "ids": [1, 2]}, {"c": -3, "ids": [1]}, {"c": 3, "ids": [1, 3]},
        {"c": 1, "ids": [1, 2]}]}}, 1.0, 9.0),
])
def test_cost_difference_estimates_match_reference(doc, mn, mx):
    """estimate_min/max_cost_diff against the reference's known answers."""
    from qiotoolkit_amd import solver as S
    m = S.parse_model(doc)
    assert m["estimate_min_cost_diff"] == pytest.approx(mn, rel=1e-15)
    if mx is not None:
        assert m["estimate_max_cost_diff"] == pytest.approx(mx, rel=1e-15)


def _largest_differencing_brute(vals):
    """Least difference by the two-largest rule on a sorted list (a second
    restatement, for the random check below)."""
    v = sorted(vals)
    while len(v) > 1:
        a = v.pop()
        b = v.pop()
        v.append(a - b)
        v.sort()
    return v[0] if v else 0.0


def test_ising_min_cost_diff_random_models():
    """Ising: min over variables of the largest-differencing remainder of its
    |c| (zeros skipped), doubled; max = 2 max_v sum |c|."""
    from qiotoolkit_amd import solver as S
    rng = np.random.default_rng(3)
    for _ in range(20):
        n = int(rng.integers(3, 9))
        terms = []
        for _t in range(int(rng.integers(n, 3 * n))):
            i, j = rng.choice(n, 2, replace=False)
            terms.append({"c": int(rng.integers(-6, 7)), "ids": [int(i), int(j)]})
        m = S.parse_model({"cost_function": {"type": "ising", "terms": terms}})
        names = m["names"]
        inc = {v: [] for v in range(len(names))}
        for t in terms:
            for x in t["ids"]:
                inc[names.index(x)].append(abs(t["c"]))
        ds = [d for d in (_largest_differencing_brute(inc[v]) for v in inc) if d != 0]
        want_min = 2 * (min(ds) if ds else 0.1)
        want_max = 2 * max(sum(inc[v]) for v in inc)
        assert m["estimate_min_cost_diff"] == pytest.approx(want_min)
        assert m["estimate_max_cost_diff"] == pytest.approx(want_max)


def test_parameter_free_plan():
    """SA/PT parameter-free: beta range from the estimates
    (sa_parameter_free.h:66-101), PT betas in geometric runs of 4, 9, 14, ...
    (pt_parameter_free.h:150-172), 5 s default time limit."""
    from qiotoolkit_amd import solver as S
    d = S.describe("paralleltempering-parameterfree.qiotoolkit", {"params": {"min_num_betas": 16}})
    b = d["all_betas"]
    assert len(b) == 16 and d["time_limit"] == 5
    lo, hi = d["beta_start"], d["beta_stop"]
    want = ([lo * (hi / lo) ** (i / 3) for i in range(4)] + [lo * (hi / lo) ** (i / 8) for i in range(9)]
            + [lo * (hi / lo) ** (i / 13) for i in range(3)])
    np.testing.assert_allclose(b, want, rtol=1e-12)
    d = S.describe("simulatedannealing-parameterfree.qiotoolkit", {"params": {"timeout": 2}})
    assert d["initial_sweeps_sample_points"] == [10, 20, 30, 50] and d["time_limit"] == 2
    assert d["min_num_restarts"] == 64


@pytest.mark.parametrize("params,msg", [
    ({"max_prob": 0}, "parameter `max_prob`: must be greater than 0"),
    ({"min_prob": 1.5}, "parameter `min_prob`: must be less-equal 1"),
    ({"initial_sweeps_sample_points": [10, 20]}, "size must be greater than 2, found 2"),
    ({"min_num_restarts": 0}, "parameter `min_num_restarts`: must be greater than 0"),
])
def test_parameter_free_configure_errors(params, msg):
    from qiotoolkit_amd import solver as S
    with pytest.raises(S.SolverError, match=re.escape(msg)):
        S.describe("simulatedannealing-parameterfree.qiotoolkit", {"params": params})


def test_c_program_compiles_and_links(tmp_path):
    """tests/c_abi/solve_fixture.c is valid C99 against include/qiotk.h and
    include/qmc.h and links against libqiotk.so and libqmc.so (-Wall
    -Werror); the GPU test runs it."""
    from tests import _c_abi
    exe = _c_abi.build(str(tmp_path))
    assert os.path.exists(exe)


def test_qio_metrics_line():
    """qiotk_metrics renders Runner::metric_to_console (app/runner.cc:640-688):
    common, solver and problem metrics between _QIO_METRICS labels, absent
    numbers as 0, the model identifier from the response."""
    from qiotoolkit_amd import solver
    resp = {"benchmark": {"solver_time_ms": 12.5, "cpu_time_ms": 3.0, "max_mem_bytes": 1024,
                          "num_terms": 10, "num_variables": 10, "locality": 2, "min_locality": 2,
                          "avg_locality": 2.0, "accumulated_dependent_vars": 20,
                          "sum_coefficient_degrees_total": 20, "max_coupling_magnitude": 1.0,
                          "min_coupling_magnitude": 1.0,
                          "extensions": [{"name": "solver", "value": {"last_step": 100}},
                                         {"name": "device", "value": {"model": "ising"}}]},
            "solutions": {"cost": -10}}
    line = solver.metrics(resp)
    assert line.startswith("_QIO_METRICS{") and line.endswith("}_QIO_METRICS")
    m = json.loads(line[len("_QIO_METRICS"):-len("_QIO_METRICS")])
    assert m["common"] == {"solver_time_ms": 12.5, "cpu_time_ms": 3.0, "max_mem_bytes": 1024,
                           "preprocessing_time_ms": 0, "postprocessing_time_ms": 0}
    assert m["solver"] == {"last_step": 100}
    assert m["problem"]["num_terms"] == 10 and m["problem"]["model"] == "ising"
    assert m["problem"]["accumulated_dependent_vars"] == 20
    with pytest.raises(solver.SolverError):
        solver.metrics({"solutions": {}})


def _tree(text):
    """The same document with an escaped key in front: the streaming reader
    takes no escapes in keys, so this one goes through the tree parser."""
    assert text.lstrip().startswith("{")
    return '{"\\u0078": 0, ' + text.lstrip()[1:]


def _both(text, streams=True):
    from qiotoolkit_amd import solver as S

    def run(t, reader):
        try:
            m = S.parse_model(t)
        except S.SolverError as e:
            return ("error", e.code, str(e))
        # which reader took it: the escaped key forces the tree parser; a
        # well-formed document without one streams
        assert m.pop("input_reader") == reader
        return ("ok", m)
    return run(text, "stream" if streams else "tree"), run(_tree(text), "tree")


@pytest.mark.parametrize("fname", ["ising1.json", "cpupubojobstest.json", "isingemptymixed.json",
                                   "puboemptymixed.json", "isingempty.json", "puboempty.json",
                                   "e2e/combined_4_local_N_16_qualitytest/combined_4_local_N_16_inst_1.json"])
def test_streaming_input_equals_tree_parse(fname):
    text = open(os.path.join(GOLD, fname)).read()
    streamed, tree = _both(text)
    assert streamed == tree


def test_streaming_input_shapes_and_errors(tmp_path):
    from qiotoolkit_amd import solver as S
    rng = np.random.default_rng(3)
    docs = []
    for kind in ("ising", "pubo"):
        terms = []
        for _ in range(300):
            k = int(rng.integers(0 if kind == "pubo" else 1, 5))
            ids = [int(x) for x in rng.choice(1000, size=k, replace=False)]
            c = float(rng.standard_normal()) if rng.random() < 0.5 else int(rng.integers(-9, 10))
            t = {"ids": ids, "c": c} if rng.random() < 0.5 else {"c": c, "ids": ids}
            if rng.random() < 0.1:
                t["note"] = {"deep": [1, 'x\\"y', None, True]}
            terms.append(t)
        names = sorted({i for t in terms for i in t["ids"]})
        ic = {str(n): (1 if kind == "pubo" else -1) if n % 2 else (0 if kind == "pubo" else 1) for n in names}
        docs.append({"cost_function": {"version": "1.1", "type": kind, "terms": terms,
                                       "initial_configuration": ic}, "metadata": {"name": "x"}})
        docs.append({"solver": [1, 2.5e-3], "cost_function": {"type": kind, "terms": terms}})
    # documents the tree parser rejects, with the reference's errors
    docs += [
        {"cost_function": {"type": "ising", "terms": []}},
        {"cost_function": {"type": "ising", "terms": [{"c": 1, "ids": [3, 3]}]}},
        {"cost_function": {"type": "ising", "version": "-7", "terms": [{"c": 1, "ids": [0]}]}},
        {"cost_function": {"type": "qubo", "terms": [{"c": 1, "ids": [0]}]}},
        {"cost_function": {"type": "ising", "terms": [{"c": 1}]}},
        {"cost_function": {"type": "ising", "terms": [{"c": 1, "ids": [0, 1.5]}]}},
        {"cost_function": {"type": "ising", "terms": [{"c": 1, "ids": [0, 1]}],
                           "initial_configuration": {"0": 1, "7": -1}}},
        {"cost_function": {"type": "ising", "terms": [{"c": 1, "ids": [0, 1]}],
                           "initial_configuration": {"0": 1, "1": 2}}},
        {"cost_function": {"type": "ising", "terms": [{"c": 1, "ids": [0, 1]}],
                           "initial_configuration": {"0": 1}}},
        {"cost_function": {"type": "ising"}},
        # the cost function at the top level: accepted, by the tree parser
        {"type": "ising", "terms": [{"c": 1, "ids": [0, 1]}]},
    ]
    for doc in docs:
        text = json.dumps(doc)
        top = "cost_function" not in doc
        streamed, tree = _both(text, streams=not top)
        assert streamed == tree, text[:200]
        # and from a file, read in 64 KiB chunks
        p = tmp_path / "in.json"
        p.write_text(text)
        try:
            got = S.parse_model_file(str(p))
            assert got.pop("input_reader") == ("tree" if top else "stream")
            got = ("ok", got)
        except S.SolverError as e:
            got = ("error", e.code, str(e))
        assert got == tree, text[:200]
    # a document larger than one file chunk, split inside numbers and keys
    big = {"cost_function": {"type": "ising", "terms": [
        {"c": float(v), "ids": [int(i), int(i) + 1]} for i, v in enumerate(rng.standard_normal(20000))]}}
    text = json.dumps(big)
    assert len(text) > 4 * 65536
    p = tmp_path / "big.json"
    p.write_text(text)
    m = S.parse_model_file(str(p))
    assert m.pop("input_reader") == "stream"
    assert m == _both(text)[1][1]
    # truncated documents fall back and fail as the tree parser fails
    for cut in (10, len(text) // 2, len(text) - 1):
        streamed, tree = _both(text[:cut])
        assert streamed[0] == tree[0] == "error" and streamed[1] == tree[1]
