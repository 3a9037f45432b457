"""ctypes binding of libqmc.so (include/qmc.h).

The library must be built (``python -m qiotoolkit_amd.build`` or
``__graft_entry__.build()``); importing this module without it raises -- there
is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import _handles

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqmc.so")

QMC_ISING, QMC_PUBO = 0, 1
QMC_COUPLING_I32, QMC_COUPLING_F32 = 0, 1
QMC_BETA_SCHEDULE, QMC_BETA_PER_REPLICA = 0, 1
QMC_INVALID_ARGUMENT = 107
# device scratch of qmc_energy_noalloc per GPU (QMC_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 16 << 20

# every symbol include/qmc.h declares
EXPORTS = [
    "qmc_last_error", "qmc_ctx_create", "qmc_ctx_destroy", "qmc_sync", "qmc_ctx_compute_units", "qmc_problem_create",
    "qmc_problem_destroy", "qmc_problem_get_info", "qmc_problem_order", "qmc_walkers_create",
    "qmc_walkers_destroy", "qmc_walkers_init", "qmc_walkers_set_states", "qmc_walkers_set_betas",
    "qmc_walkers_sweep", "qmc_walkers_pt", "qmc_walkers_read", "qmc_walkers_energy",
    "qmc_walkers_labels", "qmc_energy", "qmc_delta_all", "qmc_philox", "qmc_walkers_last_ms",
    "qmc_walkers_plan", "qmc_walkers_plan_pt", "qmc_walkers_remaining", "qmc_batch_create",
    "qmc_batch_destroy", "qmc_batch_add", "qmc_batch_remove", "qmc_batch_clear", "qmc_batch_run",
    "qmc_batch_wait", "qmc_walkers_read_replica", "qmc_batch_read", "qmc_batch_run_chunked", "qmc_walkers_resample", "qmc_walkers_count", "qmc_walkers_set_pt_mode",
    "qmc_walkers_set_lanes", "qmc_walkers_get_lanes", "qmc_walkers_pa_begin", "qmc_walkers_pa_step",
    "qmc_walkers_pa_info", "qmc_walkers_pa_end", "qmc_lds_bandwidth", "qmc_walkers_pa_strategy",
    "qmc_walkers_set_sweep_order", "qmc_problem_create_visit", "qmc_walkers_pa_observe", "qmc_walkers_pa_log",
    "qmc_walkers_pt_observe", "qmc_walkers_pt_log", "qmc_energy_noalloc",
]


class QmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("qmc error %d: %s" % (code, msg))
        self.code = code


class ProblemInfo(C.Structure):
    _fields_ = [
        ("n_vars", C.c_int32), ("n_terms", C.c_int64), ("n_incidences", C.c_int64),
        ("n_classes", C.c_int32), ("lanes", C.c_int32), ("coupling", C.c_int32),
        ("pairwise", C.c_int32), ("max_degree", C.c_int32), ("mean_degree", C.c_double),
        ("max_class", C.c_int32), ("entry_bytes", C.c_int32),
    ]


class PaInfo(C.Structure):
    _fields_ = [
        ("population", C.c_int32), ("restarts", C.c_int32), ("overflow", C.c_int32),
        ("have_best", C.c_int32), ("steps", C.c_int64), ("attempts", C.c_int64),
        ("best", C.c_double), ("beta", C.c_double), ("rho", C.c_double), ("delta_beta", C.c_double),
        ("target", C.c_int32), ("pad_", C.c_int32), ("evaluations", C.c_int64),
    ]


# problem visit orders (qmc_problem_create_visit)
VISIT_COLOR_CLASSES, VISIT_INDEX_LEVELS = 0, 1

# population annealing stepping sizes (qmc_walkers_pa_strategy)
PA_SCHEDULE, PA_FRICTION_TENSOR, PA_ENERGY_VARIANCE, PA_CONSTANT_CULLING = 0, 1, 2, 3


_lib = None


def _unload():
    """Unload libqmc.so at exit (after every handle is closed), so that its
    code-object unregistration runs while the HIP runtime is intact."""
    global _lib
    if _lib is not None:
        import _ctypes
        h, _lib = _lib._handle, None
        _ctypes.dlclose(h)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libqmc.so is not built (run __graft_entry__.build()); "
                          "the sweep engine has no CPU fallback")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    L.qmc_last_error.restype = C.c_char_p
    L.qmc_ctx_create.argtypes = [C.c_int, P(vp)]
    L.qmc_ctx_destroy.argtypes = [vp]
    L.qmc_ctx_destroy.restype = None
    L.qmc_sync.argtypes = [vp]
    L.qmc_problem_create.argtypes = [vp, C.c_int, C.c_int32, C.c_int64, P(C.c_int64), P(C.c_int32),
                                     P(C.c_double), C.c_int32, P(vp)]
    L.qmc_problem_destroy.argtypes = [vp]
    L.qmc_problem_destroy.restype = None
    L.qmc_problem_get_info.argtypes = [vp, P(ProblemInfo)]
    L.qmc_problem_order.argtypes = [vp, C.c_int32, P(C.c_int32), P(C.c_int32)]
    L.qmc_walkers_create.argtypes = [vp, C.c_int32, C.c_int32, C.c_uint32, C.c_uint32, C.c_int32,
                                     C.c_int32, P(vp)]
    L.qmc_walkers_destroy.argtypes = [vp]
    L.qmc_walkers_destroy.restype = None
    L.qmc_walkers_init.argtypes = [vp, C.c_uint32]
    L.qmc_walkers_set_states.argtypes = [vp, P(C.c_uint8)]
    L.qmc_walkers_set_betas.argtypes = [vp, P(C.c_double)]
    L.qmc_walkers_sweep.argtypes = [vp, C.c_int64, C.c_int32, C.c_int, P(C.c_double), C.c_int]
    L.qmc_walkers_pt.argtypes = [vp, C.c_int32, P(C.c_double), C.c_int64, C.c_int32, C.c_int32]
    L.qmc_walkers_read.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_uint64),
                                   P(C.c_uint8), P(C.c_uint8)]
    L.qmc_walkers_energy.argtypes = [vp, C.c_int, P(C.c_double)]
    L.qmc_walkers_labels.argtypes = [vp, P(C.c_int32)]
    L.qmc_energy.argtypes = [vp, C.c_int32, P(C.c_uint8), P(C.c_double)]
    L.qmc_energy_noalloc.argtypes = [vp, C.c_int32, P(C.c_uint8), P(C.c_double)]
    L.qmc_delta_all.argtypes = [vp, C.c_int32, P(C.c_uint8), P(C.c_double)]
    L.qmc_philox.argtypes = [vp, C.c_int32, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    L.qmc_walkers_last_ms.argtypes = [vp, P(C.c_double)]
    L.qmc_walkers_plan.argtypes = [vp, C.c_int64, C.c_int32, P(C.c_double)]
    L.qmc_walkers_plan_pt.argtypes = [vp, C.c_int32, P(C.c_double), C.c_int64, C.c_int32, C.c_int32]
    L.qmc_walkers_remaining.argtypes = [vp, P(C.c_int64)]
    L.qmc_batch_create.argtypes = [vp, P(vp)]
    L.qmc_batch_destroy.argtypes = [vp]
    L.qmc_batch_destroy.restype = None
    L.qmc_batch_add.argtypes = [vp, vp]
    L.qmc_batch_remove.argtypes = [vp, vp]
    L.qmc_batch_clear.argtypes = [vp]
    L.qmc_batch_run.argtypes = [vp, C.c_int32]
    L.qmc_batch_wait.argtypes = [vp, P(C.c_double)]
    L.qmc_batch_read.argtypes = [vp, vp, vp, vp]
    L.qmc_walkers_read_replica.argtypes = [vp, C.c_int32, vp, vp]
    L.qmc_batch_run_chunked.argtypes = [vp, C.c_int32]
    L.qmc_walkers_resample.argtypes = [vp, P(C.c_int32), C.c_int32]
    L.qmc_walkers_count.argtypes = [vp, P(C.c_int32)]
    L.qmc_walkers_set_pt_mode.argtypes = [vp, C.c_int]
    L.qmc_walkers_set_lanes.argtypes = [vp, C.c_int32]
    L.qmc_walkers_get_lanes.argtypes = [vp, P(C.c_int32)]
    L.qmc_walkers_pa_begin.argtypes = [vp, C.c_int32, C.c_double, C.c_double, P(C.c_double), C.c_int32,
                                       C.c_uint32, C.c_int32]
    L.qmc_walkers_pa_step.argtypes = [vp]
    L.qmc_walkers_pa_strategy.argtypes = [vp, C.c_int32, C.c_double, C.c_int32, C.c_int32]
    L.qmc_walkers_set_sweep_order.argtypes = [vp, C.c_int32]
    L.qmc_walkers_pa_observe.argtypes = [vp, C.c_int32]
    L.qmc_walkers_pa_log.argtypes = [vp, C.c_int32, P(C.c_double)]
    L.qmc_walkers_pt_observe.argtypes = [vp, C.c_int32, C.c_int64]
    L.qmc_walkers_pt_log.argtypes = [vp, C.c_int64, P(C.c_double)]
    L.qmc_problem_create_visit.argtypes = [vp, C.c_int, C.c_int32, C.c_int64, P(C.c_int64), P(C.c_int32),
                                           P(C.c_double), C.c_int32, C.c_int32, P(vp)]
    L.qmc_walkers_pa_info.argtypes = [vp, P(PaInfo), vp]
    L.qmc_walkers_pa_end.argtypes = [vp]
    L.qmc_lds_bandwidth.argtypes = [vp, P(C.c_double)]
    _lib = L
    _handles.on_unload(_unload)
    return L


def _check(rc):
    if rc != 0:
        raise QmcError(rc, lib().qmc_last_error().decode(errors="replace"))


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().qmc_ctx_create(device, C.byref(self.h)))
        _handles.track(self)

    def close(self):
        if self.h:
            _handles.untrack(self)
            lib().qmc_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _handles.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().qmc_sync(self.h))

    def lds_bandwidth(self):
        """Measured aggregate LDS read bandwidth in bytes/s."""
        v = C.c_double()
        _check(lib().qmc_lds_bandwidth(self.h, C.byref(v)))
        return v.value

    def philox(self, ctrs, key):
        ctrs = np.ascontiguousarray(ctrs, dtype=np.uint32).reshape(-1, 4)
        k = np.ascontiguousarray(key, dtype=np.uint32)
        out = np.zeros_like(ctrs)
        _check(lib().qmc_philox(self.h, len(ctrs), _p(ctrs, C.c_uint32), _p(k, C.c_uint32),
                                _p(out, C.c_uint32)))
        return out


class Problem:
    """A problem in internal ids (terms with sorted unique ids, no constant terms)."""

    def __init__(self, ctx, kind, n_vars, term_offsets, term_vars, coeffs, lanes=0, visit=0):
        """visit: VISIT_COLOR_CLASSES (0) or VISIT_INDEX_LEVELS (1)."""
        self.ctx = ctx
        self.kind = kind
        off = np.ascontiguousarray(term_offsets, dtype=np.int64)
        ids = np.ascontiguousarray(term_vars, dtype=np.int32)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        self.h = C.c_void_p()
        _check(lib().qmc_problem_create_visit(ctx.h, kind, n_vars, len(c), _p(off, C.c_int64),
                                              _p(ids, C.c_int32), _p(c, C.c_double), lanes, visit,
                                              C.byref(self.h)))
        _handles.track(self)
        self.info = ProblemInfo()
        _check(lib().qmc_problem_get_info(self.h, C.byref(self.info)))
        self.n = self.info.n_vars

    def close(self):
        if self.h:
            _handles.untrack(self)
            lib().qmc_problem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _handles.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def order(self, lanes=0):
        o = np.zeros(self.n, dtype=np.int32)
        co = np.zeros(self.info.n_classes + 1, dtype=np.int32)
        _check(lib().qmc_problem_order(self.h, lanes, _p(o, C.c_int32), _p(co, C.c_int32)))
        return o, co

    def energy(self, states, noalloc=False):
        """FP64 energies of plain states; noalloc: through qmc_energy_noalloc,
        which stages them in the library's fixed device scratch array."""
        s = np.ascontiguousarray(states, dtype=np.uint8).reshape(-1, self.n)
        out = np.zeros(len(s))
        fn = lib().qmc_energy_noalloc if noalloc else lib().qmc_energy
        _check(fn(self.h, len(s), _p(s, C.c_uint8), _p(out, C.c_double)))
        return out

    def delta_all(self, states):
        s = np.ascontiguousarray(states, dtype=np.uint8).reshape(-1, self.n)
        out = np.zeros((len(s), self.n))
        _check(lib().qmc_delta_all(self.h, len(s), _p(s, C.c_uint8), _p(out, C.c_double)))
        return out


class Walkers:
    def __init__(self, problem, n_replicas, replicas_per_group=8, seed=0, stream=0, first_replica=0,
                 replicas_per_seed=None):
        self.problem = problem
        self.R = n_replicas
        self.n = problem.n
        rps = replicas_per_seed if replicas_per_seed else n_replicas + first_replica
        self.h = C.c_void_p()
        _check(lib().qmc_walkers_create(problem.h, n_replicas, replicas_per_group, seed & 0xFFFFFFFF,
                                        stream, first_replica, rps, C.byref(self.h)))
        _handles.track(self)

    def close(self):
        if self.h:
            _handles.untrack(self)
            lib().qmc_walkers_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _handles.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def init(self, epoch=0):
        _check(lib().qmc_walkers_init(self.h, epoch))

    def set_states(self, states):
        s = np.ascontiguousarray(states, dtype=np.uint8).reshape(self.R, self.n)
        _check(lib().qmc_walkers_set_states(self.h, _p(s, C.c_uint8)))

    def set_betas(self, betas):
        b = np.ascontiguousarray(betas, dtype=np.float64)
        _check(lib().qmc_walkers_set_betas(self.h, _p(b, C.c_double)))

    def sweep(self, sweep0, schedule=None, n_sweeps=None, ordered=False):
        if schedule is not None:
            s = np.ascontiguousarray(schedule, dtype=np.float64)
            _check(lib().qmc_walkers_sweep(self.h, sweep0, len(s), QMC_BETA_SCHEDULE, _p(s, C.c_double),
                                           int(ordered)))
        else:
            _check(lib().qmc_walkers_sweep(self.h, sweep0, n_sweeps, QMC_BETA_PER_REPLICA, None,
                                           int(ordered)))

    def set_lanes(self, lanes):
        _check(lib().qmc_walkers_set_lanes(self.h, lanes))

    def lanes(self):
        v = C.c_int32()
        _check(lib().qmc_walkers_get_lanes(self.h, C.byref(v)))
        return v.value

    def set_pt_mode(self, mode):
        _check(lib().qmc_walkers_set_pt_mode(self.h, mode))

    def pt(self, ladder_betas, step0, n_steps, sweeps_per_replica=1):
        b = np.ascontiguousarray(ladder_betas, dtype=np.float64)
        _check(lib().qmc_walkers_pt(self.h, len(b), _p(b, C.c_double), step0, n_steps,
                                    sweeps_per_replica))

    def read(self, states=True):
        R, N = self.R, self.n
        cost, low, beta = np.zeros(R), np.zeros(R), np.zeros(R)
        acc = np.zeros(R, dtype=np.uint64)
        st = np.zeros((R, N), dtype=np.uint8) if states else None
        ls = np.zeros((R, N), dtype=np.uint8) if states else None
        _check(lib().qmc_walkers_read(self.h, _p(cost, C.c_double), _p(low, C.c_double),
                                      _p(beta, C.c_double), _p(acc, C.c_uint64),
                                      _p(st, C.c_uint8) if states else None,
                                      _p(ls, C.c_uint8) if states else None))
        return dict(cost=cost, lowest=low, beta=beta, accepted=acc, states=st, lowest_states=ls)

    def energy(self, lowest=False):
        out = np.zeros(self.R)
        _check(lib().qmc_walkers_energy(self.h, int(lowest), _p(out, C.c_double)))
        return out

    def labels(self):
        out = np.zeros(self.R, dtype=np.int32)
        _check(lib().qmc_walkers_labels(self.h, _p(out, C.c_int32)))
        return out

    def last_ms(self):
        v = C.c_double()
        _check(lib().qmc_walkers_last_ms(self.h, C.byref(v)))
        return v.value

    def plan(self, sweep0, n_sweeps, schedule=None):
        if schedule is not None:
            s = np.ascontiguousarray(schedule, dtype=np.float64)
            assert len(s) >= n_sweeps
            _check(lib().qmc_walkers_plan(self.h, sweep0, n_sweeps, _p(s, C.c_double)))
        else:
            _check(lib().qmc_walkers_plan(self.h, sweep0, n_sweeps, None))

    def plan_pt(self, ladder_betas, step0, n_steps, sweeps_per_replica=1):
        b = np.ascontiguousarray(ladder_betas, dtype=np.float64)
        _check(lib().qmc_walkers_plan_pt(self.h, len(b), _p(b, C.c_double), step0, n_steps,
                                         sweeps_per_replica))

    def resample(self, src):
        s = np.ascontiguousarray(src, dtype=np.int32)
        _check(lib().qmc_walkers_resample(self.h, _p(s, C.c_int32), len(s)))
        self.R = len(s)

    def remaining(self):
        v = C.c_int64()
        _check(lib().qmc_walkers_remaining(self.h, C.byref(v)))
        return v.value

    # population annealing with the population on the device
    def pa_begin(self, population, alpha, scale, beta_table, seed, sweeps_per_step):
        t = np.ascontiguousarray(beta_table, dtype=np.float64)
        self._pa_table = t
        _check(lib().qmc_walkers_pa_begin(self.h, population, alpha, scale, _p(t, C.c_double), len(t),
                                          seed & 0xFFFFFFFF, sweeps_per_step))

    def pa_strategy(self, strategy, constant=0.0, constant_population=False, exp_repopulation=False):
        _check(lib().qmc_walkers_pa_strategy(self.h, int(strategy), float(constant), int(bool(constant_population)),
                                             int(bool(exp_repopulation))))

    def set_sweep_order(self, ordered):
        _check(lib().qmc_walkers_set_sweep_order(self.h, int(bool(ordered))))

    def pa_observe(self, max_steps):
        _check(lib().qmc_walkers_pa_observe(self.h, int(max_steps)))

    def pa_log(self, n_steps):
        out = np.zeros((n_steps, 13))
        _check(lib().qmc_walkers_pa_log(self.h, int(n_steps), _p(out, C.c_double)))
        return out

    def pt_observe(self, ladder, max_steps):
        _check(lib().qmc_walkers_pt_observe(self.h, int(ladder), int(max_steps)))
        self._pt_obs = (ladder, max_steps)

    def pt_log(self):
        """[ladder][step][position][cost, accepted, direction, swap]"""
        ladder, steps = self._pt_obs
        out = np.zeros((self.R // ladder, steps, ladder, 4))
        _check(lib().qmc_walkers_pt_log(self.h, out.size, _p(out, C.c_double)))
        return out

    def pa_step(self):
        _check(lib().qmc_walkers_pa_step(self.h))

    def pa_info(self):
        info = PaInfo()
        best = np.zeros(max(self.n, 1), dtype=np.uint8)
        _check(lib().qmc_walkers_pa_info(self.h, C.byref(info), _p(best, C.c_uint8)))
        out = {f: getattr(info, f) for f, _ in PaInfo._fields_}
        out["best_state"] = best[:self.n]
        return out

    def pa_end(self):
        _check(lib().qmc_walkers_pa_end(self.h))
        n = C.c_int32()
        _check(lib().qmc_walkers_count(self.h, C.byref(n)))
        self.R = n.value


class Batch:
    """Work-groups of many walker sets launched together."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.h = C.c_void_p()
        self.members = []
        _check(lib().qmc_batch_create(ctx.h, C.byref(self.h)))
        _handles.track(self)

    def close(self):
        if self.h:
            _handles.untrack(self)
            lib().qmc_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _handles.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def add(self, w):
        _check(lib().qmc_batch_add(self.h, w.h))
        self.members.append(w)

    def remove(self, w):
        _check(lib().qmc_batch_remove(self.h, w.h))
        self.members = [m for m in self.members if m is not w]

    def run(self, max_sweeps):
        _check(lib().qmc_batch_run(self.h, max_sweeps))

    def run_chunked(self, sweeps_per_launch):
        _check(lib().qmc_batch_run_chunked(self.h, sweeps_per_launch))

    def wait(self):
        v = C.c_double()
        _check(lib().qmc_batch_wait(self.h, C.byref(v)))
        return v.value

    def read(self):
        """Lowest-state energies, accepted counts and lowest states of every
        member (a list of per-member dicts, in the order added)."""
        import numpy as np
        R = [m.R for m in self.members]
        N = [m.problem.n for m in self.members]
        tot_r = sum(R)
        tot_s = sum(r * n for r, n in zip(R, N))
        e = np.zeros(tot_r, dtype=np.float64)
        a = np.zeros(tot_r, dtype=np.uint64)
        s = np.zeros(max(tot_s, 1), dtype=np.uint8)
        _check(lib().qmc_batch_read(self.h, e.ctypes.data, a.ctypes.data, s.ctypes.data))
        out, ro, so = [], 0, 0
        for r, n in zip(R, N):
            out.append({"lowest": e[ro:ro + r], "accepted": a[ro:ro + r],
                        "lowest_states": s[so:so + r * n].reshape(r, n)})
            ro += r
            so += r * n
        return out
