"""Output-time driver mirroring ``run_jax_v1_parity``
(``hybrid_flood/operational/parity_runner.py:245-437``) on the MI355X backend.

Targets are ``[1e-6, 3600, 7200, ..., duration]`` (the first snapshot follows the
first step, :297-299); after each target the clock, finiteness and the step
budget are checked (:311-322) and h, speed (h > 1e-4) and node volume are
snapshotted.  The report carries the reference's field names, ledger terms
(``<term>_m3``) and the balance residuals (:333-345).  ``publish`` writes the
reference's NPZ + run_report.json layout atomically (:401-436).
"""

from __future__ import annotations

import json
import shutil
import uuid
from pathlib import Path
from time import perf_counter

import numpy as np

from .errors import ConfigurationError, SimulationError
from .solver import HipParitySolver
from .types import LEDGER_FIELDS, LoadedParityInputs


def _speed(st) -> np.ndarray:
    """parity_runner.py:237-242."""
    speed = np.zeros_like(st.h, dtype=np.float32)
    wet = st.h > 1.0e-4
    speed[wet] = np.sqrt(st.hu[wet] ** 2 + st.hv[wet] ** 2) / st.h[wet]
    return speed


def output_targets(duration_s: float, interval_s: float = 3600.0) -> np.ndarray:
    requested = np.arange(0.0, duration_s + 1.0, interval_s, dtype=np.float32)
    if requested[-1] < duration_s:
        requested = np.append(requested, np.float32(duration_s))
    return requested


def run_hip_v1_parity(inputs: LoadedParityInputs, *, duration_s: float, max_steps_per_hour: int = 20_000,
                      snapshot_interval_s: float = 3600.0, solver: HipParitySolver | None = None,
                      keep_final_state: bool = True, boundary_stage_m=None, **solver_opts) -> dict:
    if duration_s <= 0:
        raise ConfigurationError("Parity duration must be positive.")
    s = solver or HipParitySolver(inputs.surface, inputs.coupling, inputs.pipe, inputs.recharge,
                                  inputs.forcing, inputs.initial_state, boundary_stage_m=boundary_stage_m,
                                  **solver_opts)
    requested = output_targets(duration_s, snapshot_interval_s)
    snaps = _Snapshots()
    run_started = perf_counter()
    ledger = None
    for i, req in enumerate(requested):
        target = np.float32(1.0e-6) if i == 0 else req
        t0 = perf_counter()
        clock, ledger = s.advance_to(float(target), float(duration_s), max_steps_per_hour,
                                     raise_on_budget=False)
        snaps.chunk_rt.append(perf_counter() - t0)
        snaps.take(s, clock, float(target), max_steps_per_hour)
    out = _result(inputs, s, snaps, perf_counter() - run_started, ledger, duration_s, max_steps_per_hour,
                  len(requested), keep_final_state)
    if solver is None:
        s.close()
    return out


def run_hip_ensemble(members, *, duration_s: float, max_steps_per_hour: int = 20_000,
                     snapshot_interval_s: float = 3600.0, keep_final_state: bool = False,
                     **solver_opts) -> list:
    """The batched form beside ``run_hip_v1_parity``: every member (a
    ``LoadedParityInputs``) advances in one ``HipEnsemble`` to the same output
    times, each with its own CFL dt, clock, step budget and FP64 ledger.  Returns
    one result per member, in ``run_hip_v1_parity``'s format (its ``runtime_s``
    and chunk times are the batch's)."""
    from .solver import HipEnsemble
    if duration_s <= 0:
        raise ConfigurationError("Parity duration must be positive.")
    members = list(members)
    requested = output_targets(duration_s, snapshot_interval_s)
    snaps = [_Snapshots() for _ in members]
    with HipEnsemble(members, **solver_opts) as ens:
        run_started = perf_counter()
        for i, req in enumerate(requested):
            target = np.float32(1.0e-6) if i == 0 else req
            t0 = perf_counter()
            ens.advance_to(float(target), float(duration_s), max_steps_per_hour, raise_on_budget=False)
            chunk = perf_counter() - t0
            for m, sn in enumerate(snaps):
                ens.select(m)
                sn.chunk_rt.append(chunk)
                sn.take(ens, ens.clock(), float(target), max_steps_per_hour)
        runtime = perf_counter() - run_started
        out = []
        for m, (inp, sn) in enumerate(zip(members, snaps)):
            ens.select(m)
            out.append(_result(inp, ens, sn, runtime, ens.ledger(), duration_s, max_steps_per_hour,
                               len(requested), keep_final_state))
        ens.select(0)
    return out


class _Snapshots:
    """Per-run snapshot lists: depth, speed, node volume, time and step count at
    each output time, after the runner's checks (parity_runner.py:311-322)."""

    def __init__(self):
        self.depth, self.speed, self.nodes, self.times, self.steps_at, self.chunk_rt = [], [], [], [], [], []
        self.state = None

    def take(self, s, clock, target: float, max_steps_per_hour: int):
        reached = clock["time_s"]
        if reached + 1.0e-3 < target:
            raise SimulationError(f"max_steps_per_hour={max_steps_per_hour} reached only {reached:.3f}s "
                                  f"of target {target:.3f}s.")
        st = s.get_state()
        if not np.isfinite(st.h).all():
            raise SimulationError(f"Non-finite surface depth at t={reached:.3f}s.")
        if not np.isfinite(st.node_volume).all():
            raise SimulationError(f"Non-finite drainage storage at t={reached:.3f}s.")
        self.depth.append(st.h.astype(np.float32))
        self.speed.append(_speed(st))
        self.nodes.append(st.node_volume.astype(np.float32))
        self.times.append(reached)
        self.steps_at.append(int(clock["global_step"]))
        self.state = st


def _result(inputs, s, snaps: _Snapshots, runtime: float, ledger, duration_s: float, max_steps_per_hour: int,
            n_targets: int, keep_final_state: bool) -> dict:
    """The run's arrays and its report (s: the solver, with the run's member selected)."""
    area = np.asarray(inputs.surface.cell_area, np.float64)
    initial_surface = float(np.sum(np.asarray(inputs.initial_state.h, np.float64) * area))
    initial_node = float(np.sum(np.asarray(inputs.initial_state.node_volume, np.float64)))
    depth_a = np.stack(snaps.depth)
    node_a = np.stack(snaps.nodes)
    st = snaps.state
    chunk_rt, steps_at, times = snaps.chunk_rt, snaps.steps_at, snaps.times
    led = {name: float(v) for name, v in zip(LEDGER_FIELDS, ledger)}
    final_surface = float(np.sum(depth_a[-1].astype(np.float64) * area))
    final_node = float(np.sum(node_a[-1].astype(np.float64)))
    final_recharge = float(np.sum(np.asarray(st.recharge_storage, np.float64)))
    lhs = initial_surface + initial_node + led["rainfall_input"] + led["external_boundary_inflow"]
    rhs_v1 = (final_surface + final_node + led["infiltration_loss"] + led["outfall_discharge"]
              + led["unreturned_overflow"] + led["recharge_loss"] - led["boundary_net_flux"])
    rhs_complete = rhs_v1 + final_recharge
    clock = s.clock()
    steps = int(clock["global_step"])
    valid = np.asarray(inputs.surface.active, bool)
    if "building_mask" in inputs.host:
        valid = valid & ~np.asarray(inputs.host["building_mask"], bool)
    mode = s.options.surface_mode
    budget = int(n_targets * max_steps_per_hour)
    # the reference's report fields (parity_runner.py:346-388), with this backend's values
    report = {
        "implementation": "floodhip_mi355x", "model_version": "0.3.0-parity",
        "model_variant": "hip_v1_parity" if mode == "full_swe" else "hip_local_inertial",
        "topology_mode": "exact_production", "surface_step_mode": mode,
        "drainage_mode": "v1_full_dynamic_network",
        "boundary_condition": "v1_code4_outflow_only_with_historical_external_inflow",
        "backend": "hip_gfx950", "is_gpu": True, "devices": [f"hip:{int(s.options.device)}"],
        "duration_min": float(duration_s) / 60.0, "duration_s": float(duration_s), "runtime_s": runtime,
        "loop_mode": "while", "maximum_steps_per_output_interval": int(max_steps_per_hour),
        "compiled_loop_iteration_budget": budget, "executed_loop_iterations": steps,
        "inactive_tail_iterations_avoided": budget - steps,
        "chunk_runtimes_s": chunk_rt, "first_chunk_compile_and_execution_s": chunk_rt[0],
        "remaining_chunks_execution_s": float(sum(chunk_rt[1:])),
        "step_count": steps, "steps_at_snapshot": steps_at,
        "steps_per_snapshot_interval": np.diff(np.asarray([0, *steps_at], np.int64)).tolist(),
        "cfl_eval_count": int(clock["cfl_evaluations"]),
        "snapshot_count": len(times), "max_depth_m": float(np.max(depth_a)),
        "wet_triangle_count_005m": int(np.sum((depth_a[-1] >= 0.05) & valid)),
        "drainage_node_count": int(node_a.shape[1]),
        "drainage_link_count": int(np.asarray(inputs.pipe.link_from).size),
        "mapped_inlet_count": int(np.asarray(inputs.coupling.inlet_cell).size),
        "initial_surface_volume_m3": initial_surface, "initial_node_volume_m3": initial_node,
        "final_surface_volume_m3": final_surface, "final_node_volume_m3": final_node,
        "final_recharge_storage_m3": final_recharge,
        **{f"{k}_m3": v for k, v in led.items()},
        "v1_compatible_mass_residual_m3": lhs - rhs_v1,
        "v1_compatible_mass_residual_ratio": (lhs - rhs_v1) / max(abs(lhs), 1.0),
        "complete_mass_residual_m3": lhs - rhs_complete,
        "complete_mass_residual_ratio": (lhs - rhs_complete) / max(abs(lhs), 1.0),
        "validation_status": "parity_candidate_unvalidated", "reference_model": "v1_cuda_reference",
        "observationally_validated": False, "input_report": inputs.report,
    }
    out = {"snapshot_times_s": np.asarray(times, np.float32), "steps_at_snapshot": np.asarray(steps_at, np.int64),
           "depth_snapshots_m": depth_a, "speed_snapshots_mps": np.stack(snaps.speed),
           "drain_node_volume_snapshots_m3": node_a, "report": report}
    if keep_final_state:
        out["final_state"] = st
    return out


def publish(result: dict, destination, inputs: LoadedParityInputs) -> Path:
    """Write the reference's run-directory layout atomically (parity_runner.py:401-436):
    ``flood_simulation_outputs.npz`` with the reference's array names, and
    ``run_report.json``.  Host arrays the reference writes (mesh, masks, drain node
    positions) are included when the inputs carry them."""
    dest = Path(destination).resolve()
    if dest.exists():
        raise ConfigurationError(f"Parity output directory already exists: {dest}")
    partial = dest.parent / f".{dest.name}.partial-{uuid.uuid4().hex}"
    partial.mkdir(parents=True, exist_ok=False)
    host = inputs.host
    n_cells = np.asarray(inputs.surface.cell_area).shape[0]

    def first(*names):
        for n in names:
            if n in host:
                return np.asarray(host[n])
        return None

    try:
        depth = result["depth_snapshots_m"]
        st = result.get("final_state")
        arrays = dict(
            snapshot_times_min=result["snapshot_times_s"] / 60.0,
            snapshot_times_s=result["snapshot_times_s"], steps_at_snapshot=result["steps_at_snapshot"],
            depth_snapshots_m=depth, speed_snapshots_mps=result["speed_snapshots_mps"],
            drain_node_volume_snapshots_m3=result["drain_node_volume_snapshots_m3"],
            final_depth_m=depth[-1], max_depth_m=np.max(depth, axis=0),
            max_speed_mps=np.max(result["speed_snapshots_mps"], axis=0),
            final_node_volume_m3=result["drain_node_volume_snapshots_m3"][-1],
            active_surface_mask=np.asarray(inputs.surface.active, bool),
            building_mask=np.asarray(first("building_mask") if "building_mask" in host
                                     else np.zeros(n_cells, bool), bool),
            road_mask=np.asarray(first("road_mask") if "road_mask" in host else np.zeros(n_cells, bool), bool),
            terrain_elevation_m=np.asarray(first("terrain_elevation_m") if "terrain_elevation_m" in host
                                           else inputs.surface.bed, np.float32),
            drain_link_from=np.asarray(inputs.pipe.link_from, np.int32),
            drain_link_to=np.asarray(inputs.pipe.link_to, np.int32),
        )
        if st is not None:
            arrays.update(final_hu_m2ps=np.asarray(st.hu, np.float32), final_hv_m2ps=np.asarray(st.hv, np.float32),
                          final_link_flow_m3ps=np.asarray(st.link_flow, np.float32))
        for key, names, dtype in (("nodes_world", ("nodes_world", "nodes"), np.float32),
                                  ("triangles", ("triangles",), np.int32),
                                  ("centroids_world", ("centroids_world", "centroids"), np.float32),
                                  ("drain_node_x_m", ("node_x_m",), np.float32),
                                  ("drain_node_y_m", ("node_y_m",), np.float32)):
            v = first(*names)
            if v is not None:
                arrays[key] = v.astype(dtype)
        np.savez_compressed(partial / "flood_simulation_outputs.npz", **arrays)
        (partial / "run_report.json").write_text(json.dumps(result["report"], indent=2, default=_json_default),
                                                 encoding="utf-8")
        partial.rename(dest)
    except Exception:
        if partial.exists():
            shutil.rmtree(partial)
        raise
    return dest


def _json_default(value):
    if isinstance(value, (np.integer, np.floating, np.bool_)):
        return value.item()
    if isinstance(value, np.ndarray):
        return value.tolist()
    raise TypeError(f"not JSON serializable: {type(value).__name__}")
