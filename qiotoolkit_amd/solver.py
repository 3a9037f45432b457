"""Python binding of the host solver library (include/qiotk.h, libqiotk.so).

Mirrors the reference's request/response surface: a target string, a params
document ``{"params": {...}}`` and a cost function ``{"cost_function": {...}}``
(app/runner.cc:35-429), answered with ``{"benchmark": ..., "solutions": ...}``.
Errors raise :class:`SolverError` carrying the reference's error code
(utils/error_handling.h:25-52).
"""
import ctypes as C
import json
import os

from . import _handles, _qmc

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqiotk.so")

TARGETS = ("simulatedannealing.qiotoolkit", "paralleltempering.qiotoolkit", "populationannealing.cpu")

EXPORTS = ["qiotk_engine_create", "qiotk_engine_destroy", "qiotk_solve", "qiotk_model_create",
           "qiotk_model_destroy", "qiotk_solve_model", "qiotk_solve_batch", "qiotk_free",
           "qiotk_last_error", "qiotk_parse_model", "qiotk_schedule_eval", "qiotk_describe",
           "qiotk_model_create_terms", "qiotk_solve_batch_arrays", "qiotk_model_vars",
           "qiotk_tune_batch", "qiotk_server_create", "qiotk_server_submit", "qiotk_server_submit_json",
           "qiotk_server_wait", "qiotk_server_stats", "qiotk_server_destroy", "qiotk_server_last_error",
           "qiotk_model_create_file", "qiotk_solve_file", "qiotk_parse_model_file", "qiotk_observe",
           "qiotk_metrics"]


def _check_rc(rc):
    if rc:
        raise SolverError(rc, lib().qiotk_last_error().decode())


class SolverError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("error %d: %s" % (code, message))
        self.code = code
        self.message = message


_lib = None


def _unload():
    """Unload libqiotk.so at exit (after every handle is closed), so that its
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
        raise ImportError("libqiotk.so is not built (run __graft_entry__.build())")
    _qmc.lib()  # libqmc.so first
    L = C.CDLL(LIB_PATH)
    vp, P = C.c_void_p, C.POINTER
    L.qiotk_engine_create.argtypes = [C.c_int, P(vp)]
    L.qiotk_engine_destroy.argtypes = [vp]
    L.qiotk_engine_destroy.restype = None
    L.qiotk_solve.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p, P(vp)]
    L.qiotk_model_create.argtypes = [vp, C.c_char_p, P(vp)]
    L.qiotk_model_destroy.argtypes = [vp]
    L.qiotk_model_destroy.restype = None
    L.qiotk_solve_model.argtypes = [vp, C.c_char_p, C.c_char_p, P(vp)]
    L.qiotk_solve_batch.argtypes = [vp, C.c_char_p, C.c_int32, P(vp), P(C.c_char_p), P(vp),
                                    P(C.c_double)]
    L.qiotk_free.argtypes = [vp]
    L.qiotk_free.restype = None
    L.qiotk_last_error.restype = C.c_char_p
    L.qiotk_parse_model.argtypes = [C.c_char_p, P(vp)]
    L.qiotk_parse_model_file.argtypes = [C.c_char_p, P(vp)]
    L.qiotk_model_create_file.argtypes = [vp, C.c_char_p, P(vp)]
    L.qiotk_solve_file.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p, P(vp)]
    L.qiotk_schedule_eval.argtypes = [C.c_char_p, C.c_int32, P(C.c_double), P(C.c_double), P(vp)]
    L.qiotk_describe.argtypes = [C.c_char_p, C.c_char_p, P(vp)]
    L.qiotk_observe.argtypes = [C.c_char_p, C.c_char_p, P(vp)]
    L.qiotk_metrics.argtypes = [C.c_char_p, P(vp)]
    L.qiotk_model_create_terms.argtypes = [vp, C.c_int, C.c_int64, P(C.c_int64), P(C.c_int32),
                                           P(C.c_double), P(vp)]
    _lib = L
    _handles.on_unload(_unload)
    L.qiotk_solve_batch_arrays.argtypes = [vp, C.c_char_p, C.c_int32, P(vp), P(C.c_char_p), vp, vp, vp]
    L.qiotk_model_vars.argtypes = [vp, P(C.c_int32), vp]
    L.qiotk_tune_batch.argtypes = [vp, C.c_char_p, C.c_int32, P(vp), P(C.c_char_p), C.c_int32, P(vp)]
    L.qiotk_server_create.argtypes = [C.c_int, C.c_char_p, P(vp)]
    L.qiotk_server_submit.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_int, C.c_int64, P(C.c_int64),
                                      P(C.c_int32), P(C.c_double), P(C.c_int64)]
    L.qiotk_server_submit_json.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p, P(C.c_int64)]
    L.qiotk_server_wait.argtypes = [vp, C.c_int64, C.c_double, P(vp)]
    L.qiotk_server_stats.argtypes = [vp, P(vp)]
    L.qiotk_server_destroy.argtypes = [vp]
    L.qiotk_server_destroy.restype = None
    L.qiotk_server_last_error.restype = C.c_char_p
    return L


def _take(ptr):
    s = C.cast(ptr, C.c_char_p).value.decode()
    lib().qiotk_free(ptr)
    return json.loads(s)


def _doc(x):
    return x if isinstance(x, (str, bytes)) else json.dumps(x)


def _enc(x):
    x = _doc(x)
    return x.encode() if isinstance(x, str) else x


def _raise_from(resp):
    if isinstance(resp, dict) and "error" in resp:
        raise SolverError(resp["error"]["code"], resp["error"]["message"])
    return resp


class Engine:
    """One device context."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().qiotk_engine_create(device, C.byref(self.h))
        if rc:
            raise SolverError(rc, lib().qiotk_last_error().decode())
        _handles.track(self)

    def close(self):
        if self.h:
            _handles.untrack(self)
            lib().qiotk_engine_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _handles.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def solve(self, target, params, input_data):
        out = C.c_void_p()
        lib().qiotk_solve(self.h, target.encode(), _enc(params), _enc(input_data), C.byref(out))
        return _raise_from(_take(out))

    def solve_file(self, target, params, path):
        """Solve the input file `path`: a JSON document or a protobuf problem folder."""
        out = C.c_void_p()
        lib().qiotk_solve_file(self.h, target.encode(), _enc(params), os.fspath(path).encode(), C.byref(out))
        return _raise_from(_take(out))

    def model(self, input_data):
        return Model(self, input_data)

    def model_from_arrays(self, kind, offsets, ids, coeffs):
        """Model from a term CSR (kind "ising"/"pubo"), without JSON."""
        return Model(self, None, arrays=(kind, offsets, ids, coeffs))

    def solve_batch(self, target, models, params_list):
        n = len(models)
        hs = (C.c_void_p * n)(*[m.h for m in models])
        ps = (C.c_char_p * n)(*[_enc(p) for p in params_list])
        outs = (C.c_void_p * n)()
        ms = C.c_double()
        rc = lib().qiotk_solve_batch(self.h, target.encode(), n, hs, ps, outs, C.byref(ms))
        if rc:
            for o in outs:
                if o:
                    lib().qiotk_free(o)
            raise SolverError(rc, lib().qiotk_last_error().decode())
        return [_raise_from(_take(o)) for o in outs], ms.value

    def tune_batch(self, target, models, params_list, trial_sweeps=20):
        """Pick replicas per work-group and lane depth for a batch by short
        timed trials on a sample; returns {"params": {...}, "trials": [...]}."""
        n = len(models)
        hs = (C.c_void_p * n)(*[m.h for m in models])
        ps = (C.c_char_p * n)(*[_enc(p) for p in params_list])
        out = C.c_void_p()
        lib().qiotk_tune_batch(self.h, target.encode(), n, hs, ps, trial_sweeps, C.byref(out))
        return _raise_from(_take(out))

    def solve_batch_arrays(self, target, models, params_list, configs=True):
        """solve_batch with array results: (costs[n], configurations or None,
        timings {init_ms, run_ms, finalize_ms, device_ms}).  configurations is
        a list of int8 arrays by internal variable id (+1/-1 Ising, 1/0 PUBO)."""
        import numpy as np
        n = len(models)
        hs = (C.c_void_p * n)(*[m.h for m in models])
        ps = (C.c_char_p * n)(*[_enc(p) for p in params_list])
        cost = np.zeros(n, dtype=np.float64)
        sizes = [m.n_vars() for m in models]
        cfg = np.zeros(max(1, sum(sizes)), dtype=np.int8) if configs else None
        tm = np.zeros(4, dtype=np.float64)
        rc = lib().qiotk_solve_batch_arrays(self.h, target.encode(), n, hs, ps, cost.ctypes.data,
                                            cfg.ctypes.data if configs else None, tm.ctypes.data)
        if rc:
            raise SolverError(rc, lib().qiotk_last_error().decode())
        out = None
        if configs:
            out, o = [], 0
            for nv in sizes:
                out.append(cfg[o:o + nv])
                o += nv
        return cost, out, {"init_ms": tm[0], "run_ms": tm[1], "finalize_ms": tm[2], "device_ms": tm[3]}


class Server:
    """Serving loop on its own device context: submit() returns at once;
    wait() returns a request's response document once it is finished."""

    def __init__(self, device=0, **config):
        self.h = C.c_void_p()
        rc = lib().qiotk_server_create(device, json.dumps(config).encode(), C.byref(self.h))
        if rc:
            raise SolverError(rc, lib().qiotk_server_last_error().decode())
        _handles.track(self)

    def close(self):
        if self.h:
            _handles.untrack(self)
            lib().qiotk_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _handles.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def submit(self, target, params, kind, offsets, ids, coeffs):
        import numpy as np
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        rid = C.c_int64()
        rc = lib().qiotk_server_submit(self.h, target.encode(), _enc(params), 0 if kind == "ising" else 1,
                                       len(c), off.ctypes.data_as(C.POINTER(C.c_int64)),
                                       ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                       c.ctypes.data_as(C.POINTER(C.c_double)), C.byref(rid))
        if rc:
            raise SolverError(rc, lib().qiotk_server_last_error().decode())
        return rid.value

    def submit_json(self, target, params, input_data):
        rid = C.c_int64()
        rc = lib().qiotk_server_submit_json(self.h, target.encode(), _enc(params), _enc(input_data),
                                            C.byref(rid))
        if rc:
            raise SolverError(rc, lib().qiotk_server_last_error().decode())
        return rid.value

    def wait(self, request_id, timeout_ms=-1.0):
        """(status code, response document), or None while unfinished."""
        out = C.c_void_p()
        rc = lib().qiotk_server_wait(self.h, request_id, float(timeout_ms), C.byref(out))
        if not out.value:
            return None
        return rc, _take(out)

    def stats(self):
        out = C.c_void_p()
        lib().qiotk_server_stats(self.h, C.byref(out))
        return _take(out)


class Model:
    """A cost function parsed, normalised and resident on the device."""

    def __init__(self, engine, input_data, arrays=None):
        self.engine = engine
        self.h = C.c_void_p()
        if arrays is not None:
            import numpy as np
            kind, off, ids, c = arrays
            off = np.ascontiguousarray(off, dtype=np.int64)
            ids = np.ascontiguousarray(ids, dtype=np.int32)
            c = np.ascontiguousarray(c, dtype=np.float64)
            rc = lib().qiotk_model_create_terms(
                engine.h, 0 if kind == "ising" else 1, len(c), off.ctypes.data_as(C.POINTER(C.c_int64)),
                ids.ctypes.data_as(C.POINTER(C.c_int32)), c.ctypes.data_as(C.POINTER(C.c_double)),
                C.byref(self.h))
        else:
            rc = lib().qiotk_model_create(engine.h, _enc(input_data), C.byref(self.h))
        if rc:
            raise SolverError(rc, lib().qiotk_last_error().decode())
        _handles.track(self)

    def close(self):
        if self.h:
            _handles.untrack(self)
            lib().qiotk_model_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _handles.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def n_vars(self):
        # a model never changes after creation: asked once
        nv = getattr(self, "_n_vars", None)
        if nv is None:
            n = C.c_int32()
            _check_rc(lib().qiotk_model_vars(self.h, C.byref(n), None))
            nv = self._n_vars = n.value
        return nv

    def names(self):
        import numpy as np
        out = np.zeros(max(1, self.n_vars()), dtype=np.int32)
        _check_rc(lib().qiotk_model_vars(self.h, None, out.ctypes.data))
        return out[:self.n_vars()]

    def solve(self, target, params):
        out = C.c_void_p()
        lib().qiotk_solve_model(self.h, target.encode(), _enc(params), C.byref(out))
        return _raise_from(_take(out))


def parse_model(input_data):
    """graph::normalize_edges of a cost function (no device needed)."""
    out = C.c_void_p()
    lib().qiotk_parse_model(_enc(input_data), C.byref(out))
    return _raise_from(_take(out))


def parse_model_file(path):
    """The normalised model of an input file: a JSON document or a protobuf
    problem folder (no device needed)."""
    out = C.c_void_p()
    lib().qiotk_parse_model_file(os.fspath(path).encode(), C.byref(out))
    return _raise_from(_take(out))


def schedule_eval(schedule, xs):
    """Schedule::get_value at xs, plus start/stop/count/discretised values."""
    import numpy as np
    x = np.ascontiguousarray(xs, dtype=np.float64)
    y = np.zeros_like(x)
    out = C.c_void_p()
    lib().qiotk_schedule_eval(_enc(schedule), len(x), x.ctypes.data_as(C.POINTER(C.c_double)),
                              y.ctypes.data_as(C.POINTER(C.c_double)), C.byref(out))
    meta = _raise_from(_take(out))
    return y, meta


def observe(params, events):
    """The observer protocol of params["observer"] fed with events (see
    qiotk_observe); returns its render."""
    out = C.c_void_p()
    lib().qiotk_observe(_enc(params), _enc(events), C.byref(out))
    return _raise_from(_take(out))


def metrics(response):
    """The _QIO_METRICS line of a solve response (Runner::metric_to_console)."""
    out = C.c_void_p()
    rc = lib().qiotk_metrics(_enc(response), C.byref(out))
    s = C.cast(out, C.c_char_p).value.decode()
    lib().qiotk_free(out)
    if rc:
        _raise_from(json.loads(s))
    return s


def describe(target, params):
    """Solver::configure without a device: the resolved steps and betas."""
    out = C.c_void_p()
    lib().qiotk_describe(target.encode(), _enc(params), C.byref(out))
    return _raise_from(_take(out))
