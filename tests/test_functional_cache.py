"""The resident-context cache behind the functional entry points
(advance_parity_step / stable_timestep): a context is reused for the same input
arrays and rebuilt when an array is replaced or changed in place."""

import numpy as np

from floodhip.solver import _ContextCache


class _Ctx:
    built = 0

    def __init__(self):
        _Ctx.built += 1
        self._ctx = object()
        self.closed = False

    def close(self):
        self.closed = True
        self._ctx = None


def test_reuses_for_same_arrays_and_rebuilds_on_change():
    cache = _ContextCache(slots=2)
    a = np.arange(10, dtype=np.float32)
    b = np.ones(4, np.int32)
    first = cache.get(("step", (a, b), 0.85, {"precision": 32}), _Ctx)
    assert cache.get(("step", (a, b), 0.85, {"precision": 32}), _Ctx) is first
    # a scalar option is part of the key
    other = cache.get(("step", (a, b), 0.5, {"precision": 32}), _Ctx)
    assert other is not first and not first.closed
    # an array changed in place invalidates its context
    a[3] = 7.0
    fresh = cache.get(("step", (a, b), 0.85, {"precision": 32}), _Ctx)
    assert fresh is not first and first.closed
    # an equal copy is another array: another context, and the oldest is evicted
    third = cache.get(("step", (a.copy(), b), 0.85, {"precision": 32}), _Ctx)
    assert third is not fresh and other.closed
    cache.clear()
    assert fresh.closed and third.closed


def test_closed_context_is_rebuilt():
    cache = _ContextCache(slots=1)
    a = np.zeros(3, np.float32)
    c1 = cache.get((a,), _Ctx)
    c1.close()
    c2 = cache.get((a,), _Ctx)
    assert c2 is not c1 and c2._ctx is not None
