"""The allocation-free entry point fh_stable_timestep_noalloc against
fh_stable_timestep of a context holding the same state: two calls in a row on
caller-owned device arrays, the second on a smaller input, each equal bit for
bit to the context's result (a minimum of the same float32 expressions).

The caller's buffers and stream come from the HIP runtime the library itself
links (libamdhip64.so.7, already loaded with it), as a caller's would."""

import ctypes

import numpy as np
import pytest

from floodhip import _lib, synth
from floodhip.solver import HipParitySolver
from helpers import basin

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """The few HIP runtime calls a caller makes: its own buffers and stream."""

    def __init__(self):
        _lib.load()                                   # brings libamdhip64.so.7 in
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def check(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self.check(self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))))
        self.bufs.append(p)
        if a.nbytes:
            self.check(self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), H2D))
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self.check(self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), D2H))
        return out

    def stream(self):
        s = ctypes.c_void_p()
        self.check(self.rt.hipStreamCreate(ctypes.byref(s)))
        return s

    def close(self, stream=None):
        if stream is not None:
            self.rt.hipStreamDestroy(stream)
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _state(inp, seed):
    rng = np.random.default_rng(seed)
    C = inp.surface.cell_area.shape[0]
    h = rng.uniform(0.0, 2.0, C).astype(np.float32)
    h[rng.random(C) < 0.2] = 0.0
    hu = rng.normal(0.0, 0.5, C).astype(np.float32)
    hv = rng.normal(0.0, 0.5, C).astype(np.float32)
    return inp.initial_state._replace(h=h, hu=hu, hv=hv)


def _context_dt(inp, st, cfl):
    with HipParitySolver(inp.surface, forcing=inp.forcing, state=st, cfl=cfl) as s:
        return s.stable_timestep()


def _noalloc_dt(lib, st, lc, cfl, n=None):
    hip = Hip()
    stream = hip.stream()
    try:
        arrays = [hip.upload(np.asarray(a, np.float32)) for a in (st.h, st.hu, st.hv, lc)]
        out = hip.upload(np.full(1, -1.0))
        n = int(np.asarray(st.h).shape[0]) if n is None else n
        rc = lib.fh_stable_timestep_noalloc(*arrays, n, float(cfl), out, stream)
        assert rc == _lib.FH_OK, lib.fh_global_error()
        hip.check(hip.rt.hipStreamSynchronize(stream))
        return float(hip.download(out, np.zeros(1))[0])
    finally:
        hip.close(stream)


def test_noalloc_matches_context_on_two_calls_in_a_row():
    lib = _lib.load()
    big = synth.make_plain_small()                       # 20,000 cells
    small = basin(n=12, extent=120.0, code="closed")     # 288 cells
    for inp, seed, cfl in ((big, 1, 0.85), (small, 2, 0.5)):
        st = _state(inp, seed)
        st.hu[7] = np.float32(np.nan)                    # a non-finite cell is skipped by both
        ref = _context_dt(inp, st, cfl)
        got = _noalloc_dt(lib, st, inp.surface.characteristic_length, cfl)
        assert np.isfinite(ref) and ref > 0
        assert got == ref, (got, ref)


def test_noalloc_empty_and_oversized_inputs():
    lib = _lib.load()
    inp = basin(n=4, extent=40.0)
    st = _state(inp, 3)
    assert _noalloc_dt(lib, st, inp.surface.characteristic_length, 0.85, n=0) == float("inf")
    out = ctypes.c_double()
    rc = lib.fh_stable_timestep_noalloc(None, None, None, None, 2**31, 0.85, ctypes.addressof(out), None)
    assert rc == _lib.FH_ERR_VALUE and b"n_cells" in lib.fh_global_error()
