"""Python host side of the MI355X backend: the reference's solver-loop interface
over the C ABI of libfloodhip.so.

Mirrors (paths relative to HF_CUPY_JAX_COMPLETED_PROJECT/reproduction/source/src/):

* ``hybrid_flood/operational/parity_runner.py:198-234`` -- ``_advance_to_time_while``
  (same arguments and AdvanceCarry result; the carry's clock/ledger are FP64 and
  the ``*_compensation`` fields stay zero because FP64 needs no Kahan partner);
* ``hybrid_flood/jax_solver/v1_parity.py:430-469`` -- ``advance_parity_step``;
* ``hybrid_flood/jax_solver/v1_parity.py:153-158`` -- ``stable_timestep``.

Inputs are coerced to contiguous float32/int32 like the reference's CUDA
wrappers do (``gurugram_flood/kernels/triangular_swe_cuda_kernels.py:312-320``).
Errors map to the reference's exception types: ValueError for bad arguments
(``triangular_backend.py:486-489``), SimulationError for non-finite state or an
exhausted step budget (``parity_runner.py:313-322``), BackendUnavailableError
when no GPU or no built library is available (fail closed, no CPU fallback).
"""

from __future__ import annotations

import ctypes as C
from typing import NamedTuple

import numpy as np

from . import _lib
from .errors import BackendUnavailableError, SimulationError
from .types import (AdvanceCarry, CouplingTopology, MassLedger, ParityForcing, ParityState,
                    PipeTopology, RechargeTopology, SurfaceTopology, empty_coupling, empty_pipe,
                    empty_recharge)


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


class _Args:
    """Keeps converted arrays alive for the duration of one C call."""

    def __init__(self):
        self.keep = []

    def f(self, a):
        if a is None:
            return None
        a = _f32(a)
        self.keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_float))

    def i(self, a):
        if a is None:
            return None
        a = _i32(a)
        self.keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int32))

    def u8(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(np.asarray(a) > 0.5 if np.asarray(a).dtype != bool else np.asarray(a),
                                 dtype=np.uint8)
        self.keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_uint8))


BATCH_BLOCK_SIZE = 256   # HipEnsemble's default block size (plain contexts: 64, one wavefront)


class SolverOptions(NamedTuple):
    precision: int = 32
    compensated: bool = False
    fixed_dt: float | None = None
    use_graph: bool = True
    steps_per_launch: int = 32
    block_size: int = 64
    xcd_remap: bool = True
    device: int = 0
    edges_per_thread: int = 1
    pack_order: int = 0
    cell_order: int = 1
    source_classes: int = 1
    surface_mode: str = "full_swe"   # or "local_inertial" (engine.py:219-256)
    share_mesh: bool = True          # batches of one mesh store its tables once
    cfl: float = 0.85
    cfl_interval: int = 32
    reuse_factor: float = 0.8
    max_dt: float = 1.0
    dt_floor: float = 1.0e-3
    max_speed: float = 5.0


class PipeDryWeatherDiagnostics(NamedTuple):
    """triangular_backend.py:233-240 (the scalar fields), in m^3."""
    dry_weather_inflow_m3: float
    link_abs_flow_m3: float
    outfall_discharge_m3: float
    overflow_surcharge_m3: float
    final_node_volume_m3: float


# config.py:157-158 of the reference: the surface operators of the step
SURFACE_MODES = {"full_swe": 0, "local_inertial": 1}


class HipParitySolver:
    """One device-resident scenario on one GPU (one HIP stream)."""

    def __init__(self, surface: SurfaceTopology | None, coupling: CouplingTopology | None = None,
                 pipe: PipeTopology | None = None, recharge: RechargeTopology | None = None,
                 forcing: ParityForcing | None = None, state: ParityState | None = None,
                 options: SolverOptions | None = None, members: int = 1, boundary_stage_m=None, **kw):
        self.lib = _lib.load()
        opts = (options or SolverOptions())._replace(**kw)
        self.options = opts
        if self.lib.fh_device_count() <= 0:
            raise BackendUnavailableError("floodhip requires a visible HIP GPU (no CPU fallback).")
        p = _lib.FhParams()
        self.lib.fh_default_params(C.byref(p))
        p.precision = int(opts.precision)
        p.compensated = int(bool(opts.compensated))
        p.fixed_dt = float(opts.fixed_dt) if opts.fixed_dt else 0.0
        p.use_graph = int(bool(opts.use_graph))
        p.steps_per_launch = int(opts.steps_per_launch)
        p.block_size = int(opts.block_size)
        p.xcd_remap = int(bool(opts.xcd_remap))
        p.device = int(opts.device)
        p.cfl = float(opts.cfl)
        p.cfl_interval = int(opts.cfl_interval)
        p.reuse_factor = float(opts.reuse_factor)
        p.max_dt = float(opts.max_dt)
        p.dt_floor = float(opts.dt_floor)
        p.max_speed = float(opts.max_speed)
        p.edges_per_thread = int(opts.edges_per_thread)
        p.pack_order = int(opts.pack_order)
        p.cell_order = int(opts.cell_order)
        p.source_classes = int(opts.source_classes)
        if opts.surface_mode not in SURFACE_MODES:
            raise ValueError(f"surface_mode must be one of {sorted(SURFACE_MODES)}, got {opts.surface_mode!r}")
        p.surface_mode = SURFACE_MODES[opts.surface_mode]
        p.share_mesh = int(bool(opts.share_mesh))
        self._ctx = self.lib.fh_create_batch(C.byref(p), int(members))
        if not self._ctx:
            msg = self.lib.fh_global_error().decode()
            raise BackendUnavailableError(f"fh_create failed: {msg}")
        self.members = int(members)
        self._sizes = [dict(n_cells=0, n_nodes=0, n_links=0, n_recharge=0, n_receivers=0)
                       for _ in range(self.members)]
        self._member = 0
        if surface is not None:
            self.load_member(0, surface, coupling, pipe, recharge, forcing, state, boundary_stage_m)

    # per-member sizes of the selected member
    n_cells = property(lambda self: self._sizes[self._member]["n_cells"])
    n_nodes = property(lambda self: self._sizes[self._member]["n_nodes"])
    n_links = property(lambda self: self._sizes[self._member]["n_links"])
    n_recharge = property(lambda self: self._sizes[self._member]["n_recharge"])
    n_receivers = property(lambda self: self._sizes[self._member]["n_receivers"])

    def select(self, member: int):
        """Direct upload/state/clock/ledger calls to one member of a batch."""
        self._check(self.lib.fh_select_member(self._ctx, int(member)), "select_member")
        self._member = int(member)
        return self

    def load_member(self, member, surface, coupling=None, pipe=None, recharge=None, forcing=None, state=None,
                    boundary_stage_m=None):
        self.select(member)
        self.upload_surface(surface)
        if boundary_stage_m is not None:
            self.set_boundary_stage(boundary_stage_m)
        if coupling is not None and np.asarray(coupling.node_storage_area).shape[0] > 0:
            self.upload_coupling(coupling)
            self.upload_pipe(pipe if pipe is not None else empty_pipe(self.n_nodes))
        if recharge is not None and np.asarray(recharge.source_cell).shape[0] > 0:
            self.upload_recharge(recharge)
        if forcing is not None:
            self.set_forcing(forcing)
        if state is not None:
            self.set_state(state)

    # ------------------------------------------------------------------ plumbing
    def _check(self, status, what):
        if status == _lib.FH_OK:
            return
        msg = self.lib.fh_last_error(self._ctx).decode()
        if status == _lib.FH_ERR_VALUE:
            raise ValueError(f"{what}: {msg}")
        if status in (_lib.FH_ERR_SIMULATION, _lib.FH_ERR_BUDGET):
            raise SimulationError(f"{what}: {msg}")
        if status == _lib.FH_ERR_UNAVAILABLE:
            raise BackendUnavailableError(f"{what}: {msg}")
        raise RuntimeError(f"{what}: {msg}")

    def close(self):
        if getattr(self, "_ctx", None):
            self.lib.fh_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ uploads
    def upload_surface(self, s: SurfaceTopology):
        a = _Args()
        C_ = int(np.asarray(s.cell_area).shape[0])
        E_ = int(np.asarray(s.edge_left).shape[0])
        for name in ("bed", "manning_n", "cell_area", "characteristic_length", "active"):
            if np.asarray(getattr(s, name)).shape[0] != C_:
                raise ValueError(f"{name} must have one value per cell.")
        for name in ("edge_right", "edge_nx", "edge_ny", "edge_length", "boundary_code"):
            if np.asarray(getattr(s, name)).shape[0] != E_:
                raise ValueError(f"{name} must have one value per edge.")
        self._check(self.lib.fh_upload_surface(
            self._ctx, C_, E_, a.f(s.bed), a.f(s.manning_n), a.f(s.cell_area),
            a.f(s.characteristic_length), a.i(s.edge_left), a.i(s.edge_right), a.f(s.edge_nx),
            a.f(s.edge_ny), a.f(s.edge_length), a.i(s.boundary_code), a.u8(s.active)), "upload_surface")
        self._sizes[self._member]["n_cells"] = C_

    def upload_coupling(self, c: CouplingTopology):
        a = _Args()
        N = int(np.asarray(c.node_storage_area).shape[0])
        I = int(np.asarray(c.inlet_cell).shape[0])
        S = int(np.asarray(c.surcharge_cell).shape[0])
        if np.asarray(c.inlet_node).shape[0] != I:
            raise ValueError("inlet_cell and inlet_node lengths differ.")
        for name in ("node_surface_cell", "node_invert", "node_max_depth"):
            if np.asarray(getattr(c, name)).shape[0] != N:
                raise ValueError(f"{name} must have one value per node.")
        self._check(self.lib.fh_upload_coupling(
            self._ctx, N, I, S, a.i(c.inlet_cell), a.i(c.inlet_node), a.i(c.node_surface_cell),
            a.f(c.node_invert), a.f(c.node_storage_area), a.f(c.node_max_depth),
            a.f(c.inlet_capture_capacity), a.f(c.inlet_surcharge_capacity), a.i(c.surcharge_cell),
            a.i(c.surcharge_node), a.f(c.surcharge_weight), int(bool(c.surcharge_weight_normalized))),
            "upload_coupling")
        self._sizes[self._member]["n_nodes"] = N

    def upload_pipe(self, p: PipeTopology):
        a = _Args()
        L = int(np.asarray(p.link_from).shape[0])
        for name in ("link_to", "link_length", "link_manning", "link_capacity", "link_invert",
                     "link_diameter"):
            if np.asarray(getattr(p, name)).shape[0] != L:
                raise ValueError(f"{name} must have one value per link.")
        for name in ("outfall_area", "outfall_coefficient", "outfall_tailwater", "outfall_capacity"):
            if np.asarray(getattr(p, name)).shape[0] != self.n_nodes:
                raise ValueError(f"{name} must have one value per node.")
        self._check(self.lib.fh_upload_pipe(
            self._ctx, L, a.i(p.link_from), a.i(p.link_to), a.f(p.link_length), a.f(p.link_manning),
            a.f(p.link_capacity), a.f(p.link_invert), a.f(p.link_diameter), a.f(p.link_minor_loss),
            a.i(p.link_geometry_code), a.f(p.link_rect_width), a.f(p.link_rect_height),
            a.f(p.outfall_area), a.f(p.outfall_coefficient), a.f(p.outfall_tailwater),
            a.f(p.outfall_capacity)), "upload_pipe")
        self._sizes[self._member]["n_links"] = L

    def upload_recharge(self, r: RechargeTopology):
        a = _Args()
        R = int(np.asarray(r.source_cell).shape[0])
        self._check(self.lib.fh_upload_recharge(
            self._ctx, R, a.i(r.source_cell), a.f(r.storage_capacity), a.f(r.capture_capacity),
            a.f(r.recharge_capacity), a.f(r.stop_depth)), "upload_recharge")
        self._sizes[self._member]["n_recharge"] = R

    def set_boundary_stage(self, boundary_stage_m):
        """Stage values of the code-3 edges (TriangularSurfaceForcing.boundary_stage_m,
        triangular_backend.py:93-101): a scalar or one value per edge; None removes
        them (code 3 then copies the left state, as v1_parity.py does)."""
        if boundary_stage_m is None:
            self._check(self.lib.fh_set_boundary_stage(self._ctx, 0, None), "set_boundary_stage")
            return
        a = _Args()
        v = np.atleast_1d(np.asarray(boundary_stage_m, np.float32)).reshape(-1)
        self._check(self.lib.fh_set_boundary_stage(self._ctx, int(v.shape[0]), a.f(v)), "set_boundary_stage")

    def set_forcing(self, f: ParityForcing):
        a = _Args()
        nr = int(np.asarray(f.rainfall_rate_mps).shape[0])
        rcv = np.asarray(f.external_receiver_cell)
        nrcv = int(rcv.shape[0])
        times = np.asarray(f.external_times_s)
        nt = int(times.shape[0]) if nrcv else 0
        q = np.asarray(f.external_q_m3ps, dtype=np.float32).reshape(nt, nrcv) if nrcv else None
        self._check(self.lib.fh_set_forcing(
            self._ctx, nr, a.f(f.rainfall_start_s), a.f(f.rainfall_end_s), a.f(f.rainfall_rate_mps),
            a.f(f.rainfall_multiplier), a.f(f.horton_f0_mps), a.f(f.horton_fc_mps), a.f(f.horton_k_per_s),
            nt, a.f(times) if nrcv else None, nrcv, a.i(rcv) if nrcv else None,
            a.f(q) if nrcv else None), "set_forcing")
        self._sizes[self._member]["n_receivers"] = nrcv

    # ------------------------------------------------------------------ state / clock
    def set_state(self, st: ParityState):
        a = _Args()
        n = self.n_cells
        for name in ("h", "hu", "hv", "wet_time_s"):
            if np.asarray(getattr(st, name)).shape[0] != n:
                raise ValueError(f"state.{name} must have one value per cell.")
        self._check(self.lib.fh_set_state(
            self._ctx, a.f(st.h), a.f(st.hu), a.f(st.hv), a.f(st.wet_time_s),
            a.f(st.node_volume) if self.n_nodes else None,
            a.f(st.link_flow) if self.n_links else None,
            a.f(st.recharge_storage) if self.n_recharge else None), "set_state")

    def get_state(self) -> ParityState:
        n = self.n_cells
        h, hu, hv, w = (np.empty(n, np.float32) for _ in range(4))
        v = np.empty(self.n_nodes, np.float32)
        q = np.empty(self.n_links, np.float32)
        r = np.empty(self.n_recharge, np.float32)
        P = lambda x: x.ctypes.data_as(C.POINTER(C.c_float)) if x.size else None  # noqa: E731
        self._check(self.lib.fh_get_state(self._ctx, P(h), P(hu), P(hv), P(w), P(v), P(q), P(r)),
                    "get_state")
        return ParityState(h, hu, hv, w, v, q, r)

    def get_state_f64(self):
        n = self.n_cells
        h, hu, hv = (np.empty(n, np.float64) for _ in range(3))
        P = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        self._check(self.lib.fh_get_state_f64(self._ctx, P(h), P(hu), P(hv)), "get_state_f64")
        return h, hu, hv

    def get_drainage_f64(self):
        """Node volumes and link flows as float64 (the state itself when precision=64)."""
        v = np.empty(self.n_nodes, np.float64)
        q = np.empty(self.n_links, np.float64)
        P = lambda x: x.ctypes.data_as(C.POINTER(C.c_double)) if x.size else None  # noqa: E731
        self._check(self.lib.fh_get_drainage_f64(self._ctx, P(v), P(q)), "get_drainage_f64")
        return v, q

    def set_clock(self, time_s=0.0, cached_dt_s=1.0, global_step=0):
        self._check(self.lib.fh_set_clock(self._ctx, float(time_s), float(cached_dt_s), int(global_step)),
                    "set_clock")

    def clock(self) -> dict:
        c = _lib.FhClock()
        self._check(self.lib.fh_get_clock(self._ctx, C.byref(c)), "get_clock")
        return {k: getattr(c, k) for k, _ in _lib.FhClock._fields_}

    def ledger(self) -> MassLedger:
        buf = (C.c_double * _lib.LEDGER_TERMS)()
        self._check(self.lib.fh_get_ledger(self._ctx, buf), "get_ledger")
        return MassLedger(*list(buf))

    def set_ledger(self, ledger=None):
        buf = (C.c_double * _lib.LEDGER_TERMS)(*(list(ledger) if ledger is not None else [0.0] * 13))
        self._check(self.lib.fh_set_ledger(self._ctx, buf), "set_ledger")

    # ------------------------------------------------------------------ stepping
    def advance_to(self, target_time_s, final_time_s, max_steps=20_000, raise_on_budget=True):
        """Step while t < target (parity_runner.py:224-225); returns (clock dict, MassLedger)."""
        c = _lib.FhClock()
        buf = (C.c_double * _lib.LEDGER_TERMS)()
        st = self.lib.fh_advance_to(self._ctx, float(target_time_s), float(final_time_s), int(max_steps),
                                    C.byref(c), buf)
        if st == _lib.FH_ERR_BUDGET and not raise_on_budget:
            st = _lib.FH_OK
        if st == _lib.FH_ERR_BUDGET:
            raise SimulationError(
                f"max_steps_per_hour={max_steps} reached only {c.time_s:.3f}s of target {float(target_time_s):.3f}s.")
        if st == _lib.FH_ERR_SIMULATION:
            raise SimulationError(f"Non-finite surface depth or drainage storage at t={c.time_s:.3f}s.")
        self._check(st, "advance_to")
        return {k: getattr(c, k) for k, _ in _lib.FhClock._fields_}, MassLedger(*list(buf))

    def step(self, dt, rainfall_rate_mps=0.0, external_q_m3ps=None) -> MassLedger:
        """One advance_parity_step with an explicit dt; returns its StepDiagnostics."""
        if not float(dt) > 0.0:
            raise ValueError("timestep_s must be positive.")
        a = _Args()
        q = None
        if external_q_m3ps is not None and self.n_receivers:
            q = a.f(np.asarray(external_q_m3ps, np.float32).reshape(self.n_receivers))
        buf = (C.c_double * _lib.LEDGER_TERMS)()
        self._check(self.lib.fh_step(self._ctx, float(dt), float(rainfall_rate_mps), q, buf), "step")
        return MassLedger(*list(buf))

    def dry_weather_step(self, dt, node_inflow_m3ps=None, outfall_tailwater_head_m=None,
                         cap_node_overflow=True) -> "PipeDryWeatherDiagnostics":
        """step_triangular_pipe_dry_weather_network (triangular_backend.py:833-952) on the
        resident network: sewer-only flow with optional per-node inflow and tailwater
        override; no surface exchange, clock and run ledger unchanged.  The node
        volumes and link flows are read back with get_state()."""
        if not float(dt) > 0.0:
            raise ValueError("timestep_s must be positive.")
        a = _Args()
        inflow = None if node_inflow_m3ps is None else a.f(np.asarray(node_inflow_m3ps, np.float32).reshape(self.n_nodes))
        tw = None if outfall_tailwater_head_m is None else a.f(
            np.asarray(outfall_tailwater_head_m, np.float32).reshape(self.n_nodes))
        buf = (C.c_double * 5)()
        self._check(self.lib.fh_dry_weather_step(self._ctx, float(dt), inflow, tw, int(bool(cap_node_overflow)), buf),
                    "dry_weather_step")
        return PipeDryWeatherDiagnostics(*list(buf))

    def stable_timestep(self) -> float:
        out = C.c_double()
        self._check(self.lib.fh_stable_timestep(self._ctx, C.byref(out)), "stable_timestep")
        return out.value

    def profile_kernels(self, steps=10, cap=16):
        cnt = C.c_int32()
        names = C.create_string_buffer(32 * cap)
        ms = (C.c_double * cap)()
        by = (C.c_double * cap)()
        ln = (C.c_int64 * cap)()
        self._check(self.lib.fh_profile_kernels(self._ctx, int(steps), cap, C.byref(cnt), names, ms, by, ln),
                    "profile_kernels")
        out = []
        raw = names.raw
        for i in range(cnt.value):
            nm = raw[32 * i: 32 * i + 32].split(b"\0", 1)[0].decode()
            out.append({"kernel": nm, "avg_ms": ms[i], "bytes": by[i], "launches": int(ln[i])})
        return out

    def device_bytes(self) -> int:
        return int(self.lib.fh_device_bytes(self._ctx))


class HipEnsemble(HipParitySolver):
    """A batch of independent scenarios (members) advanced together on one GPU:
    every stage of a step is one launch over all members, each member keeping its
    own CFL dt, clock, forcing and FP64 ledger."""

    def __init__(self, members, options: SolverOptions | None = None, **kw):
        members = list(members)
        if options is None and len(members) > 1:
            # blocks of 256 threads unless given: a batch's members stop at their own
            # times, and every idle member's blocks are still dispatched, so fewer,
            # larger blocks cost less there (the autotuner may pick otherwise)
            kw.setdefault("block_size", BATCH_BLOCK_SIZE)
        super().__init__(None, options=options, members=len(members), **kw)
        for m, inp in enumerate(members):
            self.load_member(m, inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing,
                             inp.initial_state)
        self.select(0)

    def member_states(self):
        out = []
        for m in range(self.members):
            self.select(m)
            out.append(self.get_state())
        self.select(0)
        return out

    def member_clocks(self):
        out = []
        for m in range(self.members):
            self.select(m)
            out.append((self.clock(), self.ledger()))
        self.select(0)
        return out


# ---------------------------------------------------------------------------
# functional mirrors of the reference entry points
# ---------------------------------------------------------------------------

class _ContextCache:
    """Resident contexts of the functional entry points, so a caller that loops
    over advance_parity_step / stable_timestep uploads its topology and forcing
    once instead of on every call.

    A context is found by the identity of every input array and the value of
    every scalar option; the cache keeps those arrays alive (so their ids stay
    theirs) and a digest of their bytes, recomputed on each call, so an array
    changed in place gets a fresh context rather than a stale one.  The state is
    uploaded and read back on every call (the functions are pure).  At most
    ``slots`` contexts are kept (least recently used first out)."""

    def __init__(self, slots: int = 2):
        import threading
        from collections import OrderedDict
        self.slots = slots
        self._lock = threading.Lock()
        self._entries: "OrderedDict[tuple, tuple[list, int, HipParitySolver]]" = OrderedDict()

    @staticmethod
    def _leaves(x, out):
        if isinstance(x, np.ndarray):
            out.append(x)
            return ("a", id(x))
        if isinstance(x, (tuple, list)):
            return tuple(_ContextCache._leaves(v, out) for v in x)
        if isinstance(x, dict):
            return tuple((k, _ContextCache._leaves(x[k], out)) for k in sorted(x))
        if x is None or isinstance(x, (bool, int, float, str, np.generic)):
            return x
        arr = np.asarray(x)
        out.append(arr)
        return ("a", id(arr))

    @staticmethod
    def _digest(arrays) -> int:
        try:
            import xxhash
            h = xxhash.xxh3_64()
        except ImportError:  # pragma: no cover - xxhash ships with this image
            import hashlib
            h = hashlib.blake2b(digest_size=8)
        for a in arrays:
            h.update(f"{a.dtype.str}{a.shape}".encode())
            if a.size:
                h.update(np.ascontiguousarray(a).view(np.uint8).reshape(-1))
        return int.from_bytes(h.digest()[:8], "little")

    def get(self, inputs, build):
        """The context for ``inputs`` (any nesting of arrays, tuples and scalars),
        created by ``build()`` when none matches."""
        arrays: list = []
        key = self._leaves(inputs, arrays)
        digest = self._digest(arrays)
        with self._lock:
            hit = self._entries.pop(key, None)
            if hit is not None and hit[1] == digest and hit[2]._ctx:
                self._entries[key] = hit
                return hit[2]
            if hit is not None:
                hit[2].close()
            solver = build()
            self._entries[key] = (arrays, digest, solver)
            while len(self._entries) > self.slots:
                _, (_, _, old) = self._entries.popitem(last=False)
                old.close()
            return solver

    def clear(self):
        with self._lock:
            for _, _, s in self._entries.values():
                s.close()
            self._entries.clear()


_contexts = _ContextCache()


def clear_context_cache():
    """Release the GPU contexts kept by advance_parity_step / stable_timestep."""
    _contexts.clear()


def advance_parity_step(state: ParityState, surface: SurfaceTopology, coupling: CouplingTopology,
                        pipe: PipeTopology, recharge: RechargeTopology, *, rainfall_rate_mps,
                        rainfall_multiplier, horton_f0_mps, horton_fc_mps, horton_k_per_s,
                        external_receiver_cell, external_q_m3ps, dt, **opts):
    """v1_parity.py:430-469 on the GPU: returns (state, StepDiagnostics).

    The topology and per-cell forcing stay resident between calls with the same
    arrays (see _ContextCache); each call uploads ``state``, takes one step of
    ``dt`` and returns the new state."""
    n = np.asarray(surface.cell_area).shape[0]
    rcv = np.asarray(external_receiver_cell, np.int32).reshape(-1)
    q = np.asarray(external_q_m3ps, np.float32).reshape(-1)

    def build():
        forcing = ParityForcing(
            np.zeros(0, np.float32), np.zeros(0, np.float32), np.zeros(0, np.float32),
            np.broadcast_to(np.asarray(rainfall_multiplier, np.float32), (n,)),
            np.broadcast_to(np.asarray(horton_f0_mps, np.float32), (n,)),
            np.broadcast_to(np.asarray(horton_fc_mps, np.float32), (n,)),
            np.broadcast_to(np.asarray(horton_k_per_s, np.float32), (n,)),
            np.zeros(1 if rcv.size else 0, np.float32), rcv,
            q.reshape(1, -1) if rcv.size else np.zeros((0, 0), np.float32))
        return HipParitySolver(surface, coupling, pipe, recharge, forcing, None, **opts)

    s = _contexts.get(("step", surface, coupling, pipe, recharge, rainfall_multiplier, horton_f0_mps,
                       horton_fc_mps, horton_k_per_s, external_receiver_cell, opts), build)
    s.set_state(state)
    diag = s.step(dt, float(rainfall_rate_mps), q if rcv.size else None)
    return s.get_state(), diag


def stable_timestep(state: ParityState, topology: SurfaceTopology, cfl: float = 0.85, **opts):
    """v1_parity.py:153-158 on the GPU (the topology stays resident between calls
    with the same arrays, as in advance_parity_step)."""
    s = _contexts.get(("cfl", topology, float(cfl), opts),
                      lambda: HipParitySolver(topology, cfl=cfl, **opts))
    s.set_state(state)
    return s.stable_timestep()


class DeviceParityInputs(NamedTuple):
    """parity_runner.py:30-35."""
    surface: SurfaceTopology
    coupling: CouplingTopology
    pipe: PipeTopology
    recharge: RechargeTopology
    forcing: ParityForcing


def _advance_to_time_while(state, time_s, time_compensation_s, cached_dt_s, global_step, ledger,
                           ledger_compensation, target_time_s, final_time_s, inputs: DeviceParityInputs,
                           *, max_steps: int, solver: HipParitySolver | None = None, **opts) -> AdvanceCarry:
    """parity_runner.py:198-234: execute real timesteps until the requested output time.

    With ``solver`` given, the scenario is already resident and ``state`` may be
    None (the resident state is advanced); otherwise a context is created for the
    call.  The step budget is not an error here (the caller checks the reached
    time, parity_runner.py:313-318)."""
    own = solver is None
    s = solver or HipParitySolver(inputs.surface, inputs.coupling, inputs.pipe, inputs.recharge,
                                  inputs.forcing, None, **opts)
    try:
        if state is not None:
            s.set_state(state)
        s.set_clock(float(time_s) - float(time_compensation_s), float(cached_dt_s), int(global_step))
        s.set_ledger(ledger)
        clock, led = s.advance_to(float(target_time_s), float(final_time_s), int(max_steps),
                                  raise_on_budget=False)
        return AdvanceCarry(s.get_state(), clock["time_s"], 0.0, clock["cached_dt_s"], clock["global_step"],
                            led, MassLedger(*([0.0] * 13)))
    finally:
        if own:
            s.close()
