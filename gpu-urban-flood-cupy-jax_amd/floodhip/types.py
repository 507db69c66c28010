"""Boundary types: the reference's NamedTuples, field for field.

Field names, meaning, units and dtypes follow
``hybrid_flood/jax_solver/v1_parity.py:23-103`` (SurfaceTopology,
CouplingTopology, PipeTopology, RechargeTopology, ParityState,
StepDiagnostics), ``hybrid_flood/operational/parity_inputs.py:36-46``
(ParityForcing) and ``hybrid_flood/operational/parity_runner.py:38-61``
(MassLedger, AdvanceCarry).  Host arrays are NumPy: float32 for real fields,
int32 for indices, bool for the active mask.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

GRAVITY = 9.80665
MIN_DEPTH = 1.0e-5


class SurfaceTopology(NamedTuple):
    bed: np.ndarray
    manning_n: np.ndarray
    cell_area: np.ndarray
    characteristic_length: np.ndarray
    edge_left: np.ndarray
    edge_right: np.ndarray
    edge_nx: np.ndarray
    edge_ny: np.ndarray
    edge_length: np.ndarray
    boundary_code: np.ndarray
    active: np.ndarray


class CouplingTopology(NamedTuple):
    inlet_cell: np.ndarray
    inlet_node: np.ndarray
    node_surface_cell: np.ndarray
    node_invert: np.ndarray
    node_storage_area: np.ndarray
    node_max_depth: np.ndarray
    inlet_capture_capacity: np.ndarray
    inlet_surcharge_capacity: np.ndarray
    surcharge_cell: np.ndarray
    surcharge_node: np.ndarray
    surcharge_weight: np.ndarray
    surcharge_weight_normalized: bool


class PipeTopology(NamedTuple):
    link_from: np.ndarray
    link_to: np.ndarray
    link_length: np.ndarray
    link_area: np.ndarray
    link_radius: np.ndarray
    link_manning: np.ndarray
    link_capacity: np.ndarray
    link_invert: np.ndarray
    link_diameter: np.ndarray
    link_minor_loss: np.ndarray
    link_geometry_code: np.ndarray
    link_rect_width: np.ndarray
    link_rect_height: np.ndarray
    outfall_area: np.ndarray
    outfall_coefficient: np.ndarray
    outfall_tailwater: np.ndarray
    outfall_capacity: np.ndarray


class RechargeTopology(NamedTuple):
    source_cell: np.ndarray
    storage_capacity: np.ndarray
    capture_capacity: np.ndarray
    recharge_capacity: np.ndarray
    stop_depth: np.ndarray


class ParityState(NamedTuple):
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    wet_time_s: np.ndarray
    node_volume: np.ndarray
    link_flow: np.ndarray
    recharge_storage: np.ndarray


class ParityForcing(NamedTuple):
    rainfall_start_s: np.ndarray
    rainfall_end_s: np.ndarray
    rainfall_rate_mps: np.ndarray
    rainfall_multiplier: np.ndarray
    horton_f0_mps: np.ndarray
    horton_fc_mps: np.ndarray
    horton_k_per_s: np.ndarray
    external_times_s: np.ndarray
    external_receiver_cell: np.ndarray
    external_q_m3ps: np.ndarray


class MassLedger(NamedTuple):
    rainfall_input: float
    infiltration_loss: float
    boundary_net_flux: float
    boundary_inflow: float
    boundary_outflow: float
    external_boundary_inflow: float
    surface_capture: float
    surface_surcharge: float
    outfall_discharge: float
    overflow_surcharge: float
    unreturned_overflow: float
    recharge_capture: float
    recharge_loss: float


StepDiagnostics = MassLedger
LEDGER_FIELDS = MassLedger._fields


class AdvanceCarry(NamedTuple):
    """parity_runner.py:54-61.  The clock is float64 here and has no Kahan partner
    (time_compensation_s stays 0.0 and is kept only for signature parity)."""
    state: ParityState
    time_s: float
    time_compensation_s: float
    cached_dt_s: float
    global_step: int
    ledger: MassLedger
    ledger_compensation: MassLedger


class LoadedParityInputs(NamedTuple):
    """parity_inputs.py:49-58, without the host-asset dictionary's file paths."""
    surface: SurfaceTopology
    coupling: CouplingTopology
    pipe: PipeTopology
    recharge: RechargeTopology
    initial_state: ParityState
    forcing: ParityForcing
    host: dict
    report: dict


def empty_coupling(n_nodes: int = 0) -> CouplingTopology:
    f = np.zeros((n_nodes,), np.float32)
    i = np.zeros((0,), np.int32)
    return CouplingTopology(i, i, np.full((n_nodes,), -1, np.int32), f, f.copy(), f.copy(),
                            np.zeros((0,), np.float32), np.zeros((0,), np.float32),
                            i, i, np.zeros((0,), np.float32), False)


def empty_pipe(n_nodes: int = 0) -> PipeTopology:
    i = np.zeros((0,), np.int32)
    z = np.zeros((0,), np.float32)
    nf = np.zeros((n_nodes,), np.float32)
    return PipeTopology(i, i, z, z, z, z, z, z, z, z, i, z, z, nf, nf.copy(), nf.copy(), nf.copy())


def empty_recharge() -> RechargeTopology:
    z = np.zeros((0,), np.float32)
    return RechargeTopology(np.zeros((0,), np.int32), z, z, z, z)


def zero_state(n_cells: int, n_nodes: int = 0, n_links: int = 0, n_recharge: int = 0,
               node_volume=None) -> ParityState:
    z = np.zeros((n_cells,), np.float32)
    nv = np.zeros((n_nodes,), np.float32) if node_volume is None else np.asarray(node_volume, np.float32)
    return ParityState(z, z.copy(), z.copy(), z.copy(), nv, np.zeros((n_links,), np.float32),
                       np.zeros((n_recharge,), np.float32))
