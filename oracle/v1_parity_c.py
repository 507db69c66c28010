"""CPU oracle on all host cores: ctypes driver of oracle/v1_parity.c.

TEST INFRASTRUCTURE ONLY, like oracle/v1_parity_np.py, whose functions the C
file restates operation by operation (the reference lines are cited there):
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg use it, as the checker or the timed CPU baseline.  It covers the full
shallow-water step and the local-inertial surface mode
(``surface_mode="local_inertial"``, restating oracle/local_inertial_np.py).

``COracle`` holds one scenario's topology and forcing fields and steps a state
in place; ``advance_to_time`` is the NumPy oracle's driver (parity_runner.py:
118-158 dt policy, rain lookup, inflow interpolation, FP64 clock and ledger)
around the C step.  ``build()`` compiles the library with gcc (-O2, OpenMP,
no FP contraction) into oracle/_build/.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from oracle import v1_parity_np as orc

HERE = Path(__file__).resolve().parent
BUILD = HERE / "_build"
SRC = HERE / "v1_parity.c"
FLAGS = ["-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=c11", "-fPIC", "-shared"]


def library_path(dtype) -> Path:
    return BUILD / f"libv1oracle_{'f64' if np.dtype(dtype) == np.float64 else 'f32'}.so"


def build(verbose=False):
    """Compile the float32 and float64 oracles (always recompiles)."""
    BUILD.mkdir(exist_ok=True)
    for dtype, defs in ((np.float32, ["-DREAL=float", "-DSFX=_f32"]),
                        (np.float64, ["-DREAL=double", "-DSFX=_f64", "-DREAL_IS_DOUBLE"])):
        out = library_path(dtype)
        tmp = out.with_suffix(".so.tmp")
        cmd = ["gcc", *FLAGS, *defs, str(SRC), "-o", str(tmp), "-lm"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, out)


_P = C.c_void_p
_TOPO_FIELDS = [(n, C.c_int64) for n in ("C", "E", "I", "N", "L", "S", "Rr", "X")] + [
    (n, _P) for n in (
        "bed", "manning", "area", "lc", "edge_left", "edge_right", "code", "nx", "ny", "len", "active",
        "lptr", "rptr", "lidx", "ridx", "stage",
        "inlet_cell", "inlet_node", "node_cell", "inv", "As", "dmax", "capq", "surq", "scell", "snode", "sw")
] + [("snorm", C.c_int32)] + [
    (n, _P) for n in (
        "lfrom", "lto", "geom", "llen", "lman", "lcap", "linv", "D", "minor", "rw", "rh", "oarea", "ocoef",
        "otw", "ocap", "rcell", "rS", "rcc", "rrc", "rstop", "rainm", "f0", "fc", "hk", "rcv")
] + [("li", C.c_int32)]


class _Topo(C.Structure):
    _fields_ = _TOPO_FIELDS


class _Work(C.Structure):
    _fields_ = [(n, _P) for n in ("fl0", "fl1", "fl2", "fr0", "fr1", "fr2", "bpart", "cell_tmp", "node_tmp",
                                  "node_tmp2", "head", "inlet_a", "inlet_b", "link_a", "link_b", "entry_a",
                                  "link_pos", "sum_a", "sum_b", "cell_b", "cell_c")]


_LIBS = {}


def _load(dtype):
    key = np.dtype(dtype).name
    if key not in _LIBS:
        p = library_path(dtype)
        if not p.exists():
            build()
        lib = C.CDLL(str(p))
        sfx = "_f64" if key == "float64" else "_f32"
        real = C.c_double if key == "float64" else C.c_float
        f = getattr(lib, "oracle_step" + sfx)
        f.restype = C.c_int
        f.argtypes = [C.POINTER(_Topo), C.POINTER(_Work)] + [_P] * 7 + [C.c_double, C.c_double, _P, C.c_double,
                                                                         C.c_double, C.c_double, _P]
        g = getattr(lib, "oracle_stable_timestep" + sfx)
        g.restype = real
        g.argtypes = [C.POINTER(_Topo), _P, _P, _P, C.c_double]
        t = getattr(lib, "oracle_threads" + sfx)
        t.restype = C.c_int
        t.argtypes = [C.c_int]
        _LIBS[key] = (f, g, t)
    return _LIBS[key]


def set_threads(n: int, dtype=np.float32) -> int:
    """OpenMP threads of the oracle (0: leave as is); returns the count in use."""
    return int(_load(dtype)[2](int(n)))


def _csr(rows, n):
    """Items grouped by row, ascending item index within a row."""
    rows = np.asarray(rows, np.int64)
    order = np.argsort(rows, kind="stable")
    ptr = np.zeros(n + 1, np.int64)
    np.add.at(ptr, rows + 1, 1)
    return np.cumsum(ptr), order.astype(np.int32)


class COracle:
    """One scenario (oracle NamedTuples in ``dtype``), stepped in place on the CPU."""

    def __init__(self, surface, coupling, pipe, recharge, forcing, dtype=np.float32, boundary_stage_m=None,
                 capture_q=0.62 * 0.04, slot_wave_speed=30.0, max_speed=5.0, surface_mode="full_swe"):
        if surface_mode not in ("full_swe", "local_inertial"):
            raise ValueError(f"unknown surface_mode {surface_mode!r}")
        self.dtype = np.dtype(dtype)
        self._step, self._cfl, _ = _load(dtype)
        self.capture_q, self.slot_wave_speed, self.max_speed = capture_q, slot_wave_speed, max_speed
        self.surface, self.forcing = surface, forcing
        keep = self._keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(np.asarray(a), dtype=dt)
            if a.size == 0:
                a = np.zeros(1, dt)
            keep.append(a)
            return a.ctypes.data_as(_P)

        R, I32 = self.dtype, np.int32
        s, c, p, r, f = surface, coupling, pipe, recharge, forcing
        Cn = int(np.asarray(s.cell_area).shape[0])
        left, right = np.asarray(s.edge_left, np.int64), np.asarray(s.edge_right, np.int64)
        active = np.asarray(s.active) > 0.5
        interior = right >= 0
        ai = interior & active[left] & active[np.where(interior, right, left)]
        lptr, lidx = _csr(left, Cn)
        rrows = np.where(ai, right, -1)
        sel = np.flatnonzero(rrows >= 0)
        rptr, rord = _csr(rrows[sel], Cn)
        ridx = sel[rord].astype(np.int32)
        t = _Topo()
        t.C, t.E = Cn, int(left.shape[0])
        t.I, t.N = int(np.asarray(c.inlet_cell).shape[0]), int(np.asarray(c.node_storage_area).shape[0])
        t.L, t.S = int(np.asarray(p.link_from).shape[0]), int(np.asarray(c.surcharge_cell).shape[0])
        t.Rr, t.X = int(np.asarray(r.source_cell).shape[0]), int(np.asarray(f.external_receiver_cell).shape[0])
        t.bed, t.manning, t.area, t.lc = (arr(x, R) for x in (s.bed, s.manning_n, s.cell_area,
                                                               s.characteristic_length))
        t.edge_left, t.edge_right, t.code = arr(left, I32), arr(right, I32), arr(s.boundary_code, I32)
        t.nx, t.ny, t.len = arr(s.edge_nx, R), arr(s.edge_ny, R), arr(s.edge_length, R)
        t.active = arr(active, np.uint8)
        t.lptr, t.rptr, t.lidx, t.ridx = arr(lptr, np.int64), arr(rptr, np.int64), arr(lidx, I32), arr(ridx, I32)
        if boundary_stage_m is not None:
            t.stage = arr(np.broadcast_to(np.asarray(boundary_stage_m, R), (t.E,)), R)
        t.inlet_cell, t.inlet_node, t.node_cell = arr(c.inlet_cell, I32), arr(c.inlet_node, I32), arr(
            c.node_surface_cell, I32)
        t.inv, t.As, t.dmax = arr(c.node_invert, R), arr(c.node_storage_area, R), arr(c.node_max_depth, R)
        t.capq, t.surq = arr(c.inlet_capture_capacity, R), arr(c.inlet_surcharge_capacity, R)
        t.scell, t.snode, t.sw = arr(c.surcharge_cell, I32), arr(c.surcharge_node, I32), arr(c.surcharge_weight, R)
        t.snorm = int(bool(c.surcharge_weight_normalized))
        t.lfrom, t.lto, t.geom = arr(p.link_from, I32), arr(p.link_to, I32), arr(p.link_geometry_code, I32)
        for name, src in (("llen", p.link_length), ("lman", p.link_manning), ("lcap", p.link_capacity),
                          ("linv", p.link_invert), ("D", p.link_diameter), ("minor", p.link_minor_loss),
                          ("rw", p.link_rect_width), ("rh", p.link_rect_height), ("oarea", p.outfall_area),
                          ("ocoef", p.outfall_coefficient), ("otw", p.outfall_tailwater),
                          ("ocap", p.outfall_capacity), ("rS", r.storage_capacity), ("rcc", r.capture_capacity),
                          ("rrc", r.recharge_capacity), ("rstop", r.stop_depth),
                          ("rainm", f.rainfall_multiplier), ("f0", f.horton_f0_mps), ("fc", f.horton_fc_mps),
                          ("hk", f.horton_k_per_s)):
            setattr(t, name, arr(src, R))
        t.rcell, t.rcv = arr(r.source_cell, I32), arr(f.external_receiver_cell, I32)
        t.li = 1 if surface_mode == "local_inertial" else 0
        self.topo = t
        w = _Work()
        E, N, L = t.E, t.N, t.L
        for name in ("fl0", "fl1", "fl2", "fr0", "fr1", "fr2"):
            setattr(w, name, arr(np.zeros(E, R), R))
        w.bpart = arr(np.zeros(E), np.float64)
        w.cell_tmp = arr(np.zeros(Cn, R), R)
        for name in ("node_tmp", "node_tmp2", "head"):
            setattr(w, name, arr(np.zeros(N, R), R))
        w.inlet_a, w.inlet_b = arr(np.zeros(t.I, R), R), arr(np.zeros(t.I, R), R)
        w.link_a, w.link_b = arr(np.zeros(L, R), R), arr(np.zeros(L, R), R)
        w.entry_a = arr(np.zeros(max(t.S, t.Rr), R), R)
        w.link_pos = arr(np.zeros(max(L, N), np.uint8), np.uint8)
        w.sum_a = arr(np.zeros(max(E, Cn)), np.float64)
        w.sum_b = arr(np.zeros(max(E, Cn)), np.float64)
        w.cell_b, w.cell_c = arr(np.zeros(Cn, R), R), arr(np.zeros(Cn, R), R)
        self.work = w

    def _state_arrays(self, state):
        out = []
        for name in ("h", "hu", "hv", "wet_time_s", "node_volume", "link_flow", "recharge_storage"):
            a = np.array(getattr(state, name), dtype=self.dtype, copy=True)
            out.append(a if a.size else np.zeros(1, self.dtype))
        return out

    def step(self, state, dt, rainfall_rate_mps=0.0, external_q_m3ps=None, in_place=False):
        """advance_parity_step with an explicit dt: returns (state, diagnostics dict).
        in_place: ``state`` holds contiguous arrays of the oracle's dtype (as a
        returned state does) and is updated in place."""
        a = self._state_arrays(state) if not in_place else [
            (x if x.size else np.zeros(1, self.dtype)) for x in state]
        diag = (C.c_double * 13)()
        q = np.zeros(max(self.topo.X, 1), self.dtype) if external_q_m3ps is None else \
            np.ascontiguousarray(np.asarray(external_q_m3ps, self.dtype).reshape(-1))
        if q.size == 0:
            q = np.zeros(1, self.dtype)
        self._step(C.byref(self.topo), C.byref(self.work), *(x.ctypes.data_as(_P) for x in a),
                   float(self.dtype.type(dt)), float(rainfall_rate_mps), q.ctypes.data_as(_P), self.max_speed,
                   self.capture_q, self.slot_wave_speed, diag)
        if in_place:
            return state, dict(zip(orc.LEDGER_FIELDS, list(diag)))
        sizes = [np.asarray(getattr(state, n)).shape[0] for n in ("node_volume", "link_flow", "recharge_storage")]
        new = orc.ParityState(a[0], a[1], a[2], a[3], a[4][:sizes[0]], a[5][:sizes[1]], a[6][:sizes[2]])
        return new, dict(zip(orc.LEDGER_FIELDS, list(diag)))

    def stable_timestep(self, state, cfl=0.85):
        h, hu, hv = (np.ascontiguousarray(np.asarray(getattr(state, n)), self.dtype) for n in ("h", "hu", "hv"))
        return self.dtype.type(self._cfl(C.byref(self.topo), h.ctypes.data_as(_P), hu.ctypes.data_as(_P),
                                         hv.ctypes.data_as(_P), float(cfl)))

    def advance_to_time(self, state, clock: orc.OracleClock, ledger, target_time_s, final_time_s, *,
                        max_steps=20000, fixed_dt=None, cfl=0.85, interval=32, reuse=0.8, max_dt=1.0,
                        dt_floor=1.0e-3):
        """v1_parity_np.advance_to_time around the C step (same dt policy and lookups)."""
        F = self.dtype.type
        ledger = dict(ledger)
        steps = 0
        state = orc.ParityState(*self._state_arrays(state)[:4], *[
            np.array(getattr(state, n), dtype=self.dtype, copy=True)
            for n in ("node_volume", "link_flow", "recharge_storage")])
        while clock.time_s < target_time_s and steps < max_steps:
            if fixed_dt is not None:
                dt, evaluated = F(fixed_dt), clock.cached_dt_s
            else:
                evaluate = clock.global_step % interval == 0
                evaluated = float(self.stable_timestep(state, cfl)) if evaluate else clock.cached_dt_s
                candidate = evaluated if evaluate else float(F(F(evaluated) * F(reuse)))
                dt = F(max(min(min(candidate, max_dt), final_time_s - clock.time_s), dt_floor))
            rate = orc.rain_rate(self.forcing, clock.time_s)
            q = orc.external_q(self.forcing, clock.time_s + 0.5 * float(dt))
            state, diag = self.step(state, dt, rate, q, in_place=True)
            for k, v in diag.items():
                ledger[k] = ledger.get(k, 0.0) + v
            clock = orc.OracleClock(clock.time_s + float(dt), evaluated, clock.global_step + 1)
            steps += 1
        return state, clock, ledger
