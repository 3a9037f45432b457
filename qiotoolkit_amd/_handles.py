"""Lifetime of the native handles (contexts, problems, walker sets, batches,
engines, models, servers).

A handle is closed by its owner's ``close()``.  Handles still open when the
interpreter exits are closed by an ``atexit`` hook, newest first, while the
HIP runtime and any tool attached to it are still fully alive; the hook then
unloads the native libraries (libqiotk.so, then libqmc.so).  Unloading runs
each library's code-object unregistration (the ``__hip_module_dtor`` that
hipcc emits and registers with ``atexit`` when the library is loaded) at that
point.  Left to the C runtime's exit handlers it would run after those of
anything that attached to the HIP runtime later -- a profiler's tool
finalisation among them -- and call into a runtime whose interception
tables are already gone.  Finalisers
(``__del__``) close a handle only during normal execution: once the
interpreter is finalising they do nothing, because destroying a context then
would call hipStreamSynchronize / hipStreamDestroy / hipHostFree after the
runtime (or a profiler wrapping it) may already have shut down.
"""
import atexit
import itertools
import sys
import weakref

_live = {}
_seq = itertools.count()
_unloaders = []


def track(obj):
    """Register an object with a ``close()`` method that owns a native handle."""
    _live[id(obj)] = (next(_seq), weakref.ref(obj))


def untrack(obj):
    _live.pop(id(obj), None)


def on_unload(fn):
    """Run fn() at exit after every handle is closed (newest first)."""
    _unloaders.append(fn)


def finalizing():
    return sys.is_finalizing()


def close_all():
    """Close every handle still open, newest first (servers and walker sets
    before the models and contexts they were created on)."""
    for _, ref in sorted(_live.values(), key=lambda t: -t[0]):
        obj = ref()
        if obj is None:
            continue
        try:
            obj.close()
        except Exception:
            pass
    _live.clear()
    while _unloaders:
        try:
            _unloaders.pop()()
        except Exception:
            pass


atexit.register(close_all)
