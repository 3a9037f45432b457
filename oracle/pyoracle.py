"""ctypes wrapper over the CPU oracle (oracle/oracle.cc) -- test infrastructure.

Imported only by tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline
leg, always as the checker.  The product package never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

RNG_REFERENCE, RNG_COUNTER = 0, 1
ARITH_F64, ARITH_F32 = 0, 1
ISING, PUBO = 0, 1


class Opts(C.Structure):
    _fields_ = [
        ("rng_mode", C.c_int),
        ("arith", C.c_int),
        ("n_order", C.c_int),
        ("order", C.POINTER(C.c_int32)),
        ("seed", C.c_uint32),
        ("stream", C.c_uint32),
        ("threads", C.c_int),
        ("steps_per_tick", C.c_int64),
    ]


_libs = {}


def build(variant=""):
    import subprocess
    target = "liboracle%s.so" % variant
    subprocess.run(["make", "-s", "-C", _HERE, target], check=True)
    return os.path.join(_HERE, target)


def lib(variant=""):
    """Load liboracle{variant}.so ('' = IEEE parity build, '_fast' = release flags)."""
    if variant in _libs:
        return _libs[variant]
    path = os.path.join(_HERE, "liboracle%s.so" % variant)
    if not os.path.exists(path):
        build(variant)
    L = C.CDLL(path)
    P = C.POINTER
    L.orc_pcg_uint32.argtypes = [C.c_int, C.c_uint32, C.c_int, P(C.c_uint32)]
    L.orc_pcg_uniform.argtypes = [C.c_int, C.c_uint32, C.c_int, P(C.c_double)]
    L.orc_pcg_fork_uniform.argtypes = [C.c_uint32, C.c_int, C.c_int, P(C.c_double)]
    L.orc_pcg_poisson.argtypes = [C.c_uint32, C.c_double, C.c_int, P(C.c_uint32)]
    L.orc_poisson_counter.argtypes = [C.c_double, C.c_uint32, C.c_uint32, C.c_int, P(C.c_uint32)]
    L.orc_pa_step_counter.argtypes = [C.c_int, P(C.c_double), P(C.c_double), P(C.c_int32), C.c_int64,
                                      C.c_int64, P(C.c_double), C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_uint32, C.c_int, C.c_double, P(C.c_int32),
                                      P(C.c_int32), P(C.c_int32), P(C.c_double), P(C.c_double), C.c_int,
                                      C.c_double, C.c_int, C.c_int, C.c_int]
    L.orc_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    L.orc_uniform_ln.argtypes = [P(C.c_double)]
    L.orc_accept.argtypes = [C.c_double, C.c_double, P(C.c_uint32), P(C.c_int)]
    L.orc_accept.restype = C.c_int
    L.orc_schedule_value.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_schedule_value.restype = C.c_double
    L.orc_schedule_discretized.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, P(C.c_double)]
    L.orc_problem_new.argtypes = [C.c_int, C.c_int64, P(C.c_double), P(C.c_int64), P(C.c_int32),
                                  C.c_char_p, C.c_int]
    L.orc_problem_new.restype = C.c_void_p
    L.orc_problem_free.argtypes = [C.c_void_p]
    L.orc_problem_dims.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64), P(C.c_int64),
                                   P(C.c_double), P(C.c_int)]
    L.orc_problem_names.argtypes = [C.c_void_p, P(C.c_int32)]
    L.orc_problem_edges.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int32), P(C.c_double)]
    L.orc_cost.argtypes = [C.c_void_p, P(C.c_uint8)]
    L.orc_cost.restype = C.c_double
    L.orc_delta.argtypes = [C.c_void_p, P(C.c_uint8), C.c_int64, C.c_int]
    L.orc_delta.restype = C.c_double
    L.orc_check_classes.argtypes = [C.c_void_p, P(C.c_int32), C.c_int, P(C.c_int32)]
    L.orc_check_classes.restype = C.c_int
    L.orc_initial_state.argtypes = [C.c_void_p, P(Opts), C.c_int32, C.c_uint32, P(C.c_uint8)]
    L.orc_run_sweeps.argtypes = [C.c_void_p, P(Opts), C.c_int, P(C.c_int32), C.c_int64, C.c_int,
                                 P(C.c_double), P(C.c_uint8), P(C.c_double), P(C.c_double),
                                 P(C.c_uint8), P(C.c_uint64)]
    L.orc_sa.argtypes = [C.c_void_p, P(Opts), C.c_int, C.c_double, C.c_double, C.c_int, C.c_int64,
                         C.c_int, P(C.c_double), P(C.c_uint8), P(C.c_double), P(C.c_uint8),
                         P(C.c_int64)]
    L.orc_pt.argtypes = [C.c_void_p, P(Opts), P(C.c_double), C.c_int, C.c_int, C.c_int64, C.c_int,
                         P(C.c_double), P(C.c_double), P(C.c_uint8), P(C.c_double), P(C.c_uint8),
                         P(C.c_double), P(C.c_int64), P(C.c_double), P(C.c_double), P(C.c_double)]
    L.orc_pa.argtypes = [C.c_void_p, P(Opts), C.c_int, C.c_double, C.c_double, C.c_int64, C.c_int,
                         C.c_int, C.c_double, C.c_int, P(C.c_double), P(C.c_uint8), P(C.c_int64),
                         P(C.c_int64), C.c_int, C.c_double, C.c_int]
    L.orc_toy_sa.argtypes = [C.c_uint32, C.c_int64, C.c_int, C.c_double, C.c_double, P(C.c_double),
                             P(C.c_double), C.c_int]
    L.orc_toy_pt.argtypes = [C.c_uint32, C.c_int64, C.c_int, P(C.c_double), C.c_int] + [P(C.c_double)] * 7 + [C.c_int]
    L.orc_selector_draw.argtypes = [C.c_int, P(C.c_double), C.c_int, P(C.c_double), P(C.c_int64)]
    L.orc_adaptive_delta_beta.argtypes = [C.c_int, C.c_double, C.c_int, P(C.c_double), P(C.c_double),
                                          C.c_int, C.c_double, C.c_int]
    L.orc_adaptive_delta_beta.restype = C.c_double
    L.orc_culling_estimate.argtypes = [C.c_int, P(C.c_double), C.c_double, C.c_int, C.c_double, C.c_double]
    L.orc_culling_estimate.restype = C.c_double
    _libs[variant] = L
    return L


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def flatten_terms(terms):
    """[{'c': c, 'ids': [...]}, ...] -> (c float64, offsets int64, ids int32)."""
    c = np.array([float(t["c"]) for t in terms], dtype=np.float64)
    lens = [len(t["ids"]) for t in terms]
    offsets = np.zeros(len(terms) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(lens)
    ids = np.array([i for t in terms for i in t["ids"]], dtype=np.int32)
    return c, offsets, ids


def make_opts(rng_mode=RNG_REFERENCE, seed=0, arith=ARITH_F64, order=None, stream=0, threads=0,
              steps_per_tick=10):
    o = Opts()
    o.rng_mode = rng_mode
    o.arith = arith
    o.seed = seed & 0xFFFFFFFF
    o.stream = stream
    o.threads = threads
    o.steps_per_tick = steps_per_tick
    if order is not None:
        arr = np.ascontiguousarray(order, dtype=np.int32)
        o._order_keep = arr
        o.n_order = len(arr)
        o.order = _p(arr, C.c_int32)
    else:
        o.n_order = 0
    return o


class Problem:
    """Normalised problem in the oracle (graph::normalize_edges semantics)."""

    def __init__(self, kind, terms, variant=""):
        """terms: [{'c': c, 'ids': [...]}, ...] or a term CSR (offsets, ids, coeffs)."""
        self.L = lib(variant)
        self.kind = ISING if kind in ("ising", ISING) else PUBO
        if isinstance(terms, tuple):
            off, ids, c = (np.ascontiguousarray(terms[0], dtype=np.int64),
                           np.ascontiguousarray(terms[1], dtype=np.int32),
                           np.ascontiguousarray(terms[2], dtype=np.float64))
        else:
            c, off, ids = flatten_terms(terms)
        err = C.create_string_buffer(256)
        h = self.L.orc_problem_new(self.kind, len(c), _p(c, C.c_double), _p(off, C.c_int64),
                                   _p(ids, C.c_int32) if len(ids) else None, err, 256)
        if not h:
            raise ValueError(err.value.decode())
        self.h = h
        n, m, inc = C.c_int64(), C.c_int64(), C.c_int64()
        cc, ml = C.c_double(), C.c_int()
        self.L.orc_problem_dims(h, C.byref(n), C.byref(m), C.byref(inc), C.byref(cc), C.byref(ml))
        self.n, self.m, self.n_inc = n.value, m.value, inc.value
        self.const_cost, self.max_locality = cc.value, ml.value

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_problem_free(self.h)
            self.h = None

    def names(self):
        out = np.zeros(self.n, dtype=np.int32)
        self.L.orc_problem_names(self.h, _p(out, C.c_int32))
        return out

    def edges(self):
        off = np.zeros(self.m + 1, dtype=np.int64)
        ids = np.zeros(self.n_inc, dtype=np.int32)
        c = np.zeros(self.m, dtype=np.float64)
        self.L.orc_problem_edges(self.h, _p(off, C.c_int64), _p(ids, C.c_int32), _p(c, C.c_double))
        return off, ids, c

    def cost(self, spins):
        s = np.ascontiguousarray(spins, dtype=np.uint8)
        return self.L.orc_cost(self.h, _p(s, C.c_uint8))

    def delta(self, spins, i, arith=ARITH_F64):
        s = np.ascontiguousarray(spins, dtype=np.uint8)
        return self.L.orc_delta(self.h, _p(s, C.c_uint8), i, arith)

    def check_classes(self, order, class_offsets):
        o = np.ascontiguousarray(order, dtype=np.int32)
        co = np.ascontiguousarray(class_offsets, dtype=np.int32)
        return bool(self.L.orc_check_classes(self.h, _p(o, C.c_int32), len(co) - 1, _p(co, C.c_int32)))

    def initial_state(self, opts, replica, epoch=0):
        out = np.zeros(self.n, dtype=np.uint8)
        self.L.orc_initial_state(self.h, C.byref(opts), replica, epoch, _p(out, C.c_uint8))
        return out

    def run_sweeps(self, opts, states, costs, lowest, lowest_states, replica_ids, sweep0, betas):
        R = states.shape[0]
        betas = np.ascontiguousarray(betas, dtype=np.float64).reshape(-1, R)
        rid = np.ascontiguousarray(replica_ids, dtype=np.int32)
        acc = np.zeros(R, dtype=np.uint64)
        rc = self.L.orc_run_sweeps(self.h, C.byref(opts), R, _p(rid, C.c_int32), sweep0, betas.shape[0],
                                   _p(betas, C.c_double), _p(states, C.c_uint8), _p(costs, C.c_double),
                                   _p(lowest, C.c_double), _p(lowest_states, C.c_uint8),
                                   _p(acc, C.c_uint64))
        if rc != 0:
            raise RuntimeError("orc_run_sweeps failed")
        return acc

    def sa(self, opts, beta_start, beta_stop, sweeps, restarts, sched_type=0, use_inverse=True):
        R, N = restarts, self.n
        lowest = np.zeros(R)
        lstates = np.zeros((R, N), dtype=np.uint8)
        fc = np.zeros(R)
        fs = np.zeros((R, N), dtype=np.uint8)
        steps = C.c_int64()
        rc = self.L.orc_sa(self.h, C.byref(opts), sched_type, beta_start, beta_stop, int(use_inverse),
                           sweeps, R, _p(lowest, C.c_double), _p(lstates, C.c_uint8), _p(fc, C.c_double),
                           _p(fs, C.c_uint8), C.byref(steps))
        if rc != 0:
            raise ValueError("orc_sa rejected the parameters")
        return dict(lowest=lowest, lowest_states=lstates, final_costs=fc, final_states=fs,
                    steps=steps.value)

    def pt(self, opts, values, use_inverse, sweeps, sweeps_per_replica=1):
        vals = np.ascontiguousarray(values, dtype=np.float64)
        R, N = len(vals), self.n
        sb = np.zeros(R)
        lowest = np.zeros(R)
        lstates = np.zeros((R, N), dtype=np.uint8)
        fc = np.zeros(R)
        fs = np.zeros((R, N), dtype=np.uint8)
        sr = np.zeros(R)
        ar, ac, ad = np.zeros(R), np.zeros(R), np.zeros(R)
        steps = C.c_int64()
        rc = self.L.orc_pt(self.h, C.byref(opts), _p(vals, C.c_double), R, int(use_inverse), sweeps,
                           sweeps_per_replica, _p(sb, C.c_double), _p(lowest, C.c_double),
                           _p(lstates, C.c_uint8), _p(fc, C.c_double), _p(fs, C.c_uint8),
                           _p(sr, C.c_double), C.byref(steps), _p(ar, C.c_double), _p(ac, C.c_double),
                           _p(ad, C.c_double))
        if rc != 0:
            raise ValueError("orc_pt rejected the parameters")
        return dict(slot_beta=sb, lowest=lowest, lowest_states=lstates, final_costs=fc,
                    final_states=fs, swap_rate=sr, steps=steps.value, acc_rate=ar, avg_cost=ac, avg_dir=ad)

    def pa(self, opts, beta_start, beta_stop, sweeps, population, sweeps_per_replica=1, alpha=2.0,
           geometric=False, constant_population=False, strategy=0, strategy_const=0.0, exp_repopulation=False):
        """strategy: 0 schedule, 1 friction tensor, 2 energy variance, 3 constant
        culling (strategy_const = its constant)."""
        bc = C.c_double()
        bs = np.zeros(self.n, dtype=np.uint8)
        fp, nr = C.c_int64(), C.c_int64()
        rc = self.L.orc_pa(self.h, C.byref(opts), int(geometric), beta_start, beta_stop, sweeps,
                           population, sweeps_per_replica, alpha, int(constant_population),
                           C.byref(bc), _p(bs, C.c_uint8), C.byref(fp), C.byref(nr), int(strategy),
                           float(strategy_const), int(exp_repopulation))
        if rc != 0:
            raise ValueError("orc_pa rejected the parameters (rc=%d)" % rc)
        return dict(best=bc.value, best_state=bs, final_population=fp.value, restarts=nr.value)


def pcg_uniform(n, seed=None):
    out = np.zeros(n)
    lib().orc_pcg_uniform(0 if seed is None else 1, 0 if seed is None else seed, n, _p(out, C.c_double))
    return out


def pcg_fork_uniform(seed, n_forks, n):
    out = np.zeros((n_forks + 1) * n)
    lib().orc_pcg_fork_uniform(seed, n_forks, n, _p(out, C.c_double))
    return out.reshape(n_forks + 1, n)


def poisson_counter(lam, seed, step, n):
    """Counter-mode Poisson draws for citizens 0..n-1 of a step."""
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_poisson_counter(lam, seed & 0xFFFFFFFF, step, n, _p(out, C.c_uint32))
    return out


def pa_step_counter(cost, lowest, family, step, base_step, beta_table, target, cap, alpha, scale, seed,
                    have_best, best, cost_before=None, strategy=0, strategy_const=0.0,
                    constant_population=False, sweeps_per_step=1, exp_repopulation=False):
    """One counter-mode population annealing step's decisions (orc_pa_step_counter)."""
    cb = None if cost_before is None else np.ascontiguousarray(cost_before, dtype=np.float64)
    cost = np.ascontiguousarray(cost, dtype=np.float64)
    lowest = np.ascontiguousarray(lowest, dtype=np.float64)
    family = np.ascontiguousarray(family, dtype=np.int32)
    table = np.ascontiguousarray(beta_table, dtype=np.float64)
    R = len(cost)
    src = np.zeros(max(cap, target), dtype=np.int32)
    fam = np.zeros(max(cap, target), dtype=np.int32)
    iout = np.zeros(6, dtype=np.int32)
    dout = np.zeros(3)
    rc = lib().orc_pa_step_counter(R, _p(cost, C.c_double), _p(lowest, C.c_double), _p(family, C.c_int32),
                                   step, base_step, _p(table, C.c_double), len(table), target, cap, alpha,
                                   scale, seed & 0xFFFFFFFF, int(have_best), best, _p(src, C.c_int32),
                                   _p(fam, C.c_int32), _p(iout, C.c_int32), _p(dout, C.c_double),
                                   None if cb is None else _p(cb, C.c_double), int(strategy),
                                   float(strategy_const), int(constant_population), int(sweeps_per_step),
                                   int(exp_repopulation))
    if rc != 0:
        raise ValueError("orc_pa_step_counter failed (rc=%d)" % rc)
    Rf = int(iout[0])
    return dict(src=src[:Rf], family=fam[:Rf], R=Rf, R_resampled=int(iout[1]), moved=bool(iout[2]),
                restart=bool(iout[3]), best_src=int(iout[4]), target=int(iout[5]), rho=float(dout[0]),
                dbeta=float(dout[1]),
                best=float(dout[2]))


def toy_sa(seed, step_limit, restarts, beta_start, beta_stop):
    """SA on the reference's toy model: the lowest-cost milestones [(step, cost)]."""
    st, va = np.zeros(1000), np.zeros(1000)
    n = lib().orc_toy_sa(seed, step_limit, restarts, beta_start, beta_stop, _p(st, C.c_double),
                         _p(va, C.c_double), 1000)
    if n < 0:
        raise ValueError("orc_toy_sa rejected the parameters")
    return list(zip(st[:n], va[:n]))


def toy_pt(seed, step_limit, temps, sweeps_per_replica=1):
    """PT on the reference's toy model: per-replica step averages and milestones."""
    n = len(temps)
    t = np.ascontiguousarray(temps, dtype=np.float64)
    outs = [np.zeros(n) for _ in range(5)]
    st, va = np.zeros(1000), np.zeros(1000)
    k = lib().orc_toy_pt(seed, step_limit, n, _p(t, C.c_double), sweeps_per_replica,
                         *[_p(o, C.c_double) for o in outs], _p(st, C.c_double), _p(va, C.c_double), 1000)
    if k < 0:
        raise ValueError("orc_toy_pt rejected the parameters")
    names = ("T", "acc_rate", "avg_cost", "avg_dir", "swap_rate")
    return dict(zip(names, outs)), list(zip(st[:k], va[:k]))


def selector_draw(weights, uniforms):
    """utils::RandomSelector picks without replacement (orc_selector_draw)."""
    w = np.ascontiguousarray(weights, dtype=np.float64)
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    out = np.zeros(len(u), dtype=np.int64)
    if lib().orc_selector_draw(len(w), _p(w, C.c_double), len(u), _p(u, C.c_double), _p(out, C.c_int64)) != 0:
        raise ValueError("orc_selector_draw rejected the sizes")
    return out


def adaptive_delta_beta(strategy, c, before, after, target, scale=1.0, sweeps=1):
    b = np.ascontiguousarray(before, dtype=np.float64)
    a = np.ascontiguousarray(after, dtype=np.float64)
    return lib().orc_adaptive_delta_beta(strategy, c, len(a), _p(b, C.c_double), _p(a, C.c_double), target,
                                         scale, sweeps)


def culling_estimate(shifted, epsilon0, target, scale, db):
    x = np.ascontiguousarray(shifted, dtype=np.float64)
    return lib().orc_culling_estimate(len(x), _p(x, C.c_double), epsilon0, target, scale, db)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


def uniform_ln():
    out = np.zeros(1024)
    lib().orc_uniform_ln(_p(out, C.c_double))
    return out.reshape(4, 256)


def accept(cost_diff, beta, draws):
    d = np.zeros(4, dtype=np.uint32)
    d[: len(draws)] = draws
    used = C.c_int()
    a = lib().orc_accept(cost_diff, beta, _p(d, C.c_uint32), C.byref(used))
    return bool(a), used.value
