"""CPU oracle: NumPy restatement of the reference's JAX V1-parity solver step.

TEST INFRASTRUCTURE ONLY.  Nothing on the product path imports this module:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker / the timed CPU baseline.

Every function restates one function of the reference (paths relative to
``HF_CUPY_JAX_COMPLETED_PROJECT/reproduction/source/src/``):

* ``hybrid_flood/jax_solver/v1_parity.py`` (the semantic oracle the reference
  itself uses for cross-backend parity), and
* ``hybrid_flood/operational/parity_runner.py`` (dt policy, forcing lookup,
  compensated clock and ledger).

The arithmetic follows the reference expression by expression, in the same
operation order, so that a float32 run rounds like the reference's float32
JAX run does.  ``dtype`` selects float32 (the reference's precision) or
float64 (the FP64 oracle that precision results are reported against).
Scatter-adds use ``np.add.at`` (sequential, index order), matching the order
the reference's CPU scatter uses.

Pinning (see tests/test_oracle_*.py): every function here is checked against
closed forms, conservation laws or brute force -- lake at rest, mirror
symmetry, volume conservation in a closed basin, rain-only h = r t, the
Horton closed form, the Stoker dam-break solution, a hand-computed
two-triangle flux, circular-segment geometry identities, donor-limiting
invariants and the ledger balance identity.  The reference ships no golden
arrays for this path (its data files are absent), so no function here is
pinned by reference-held vectors.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

GRAVITY = 9.80665            # v1_parity.py:14
MIN_DEPTH = 1.0e-5           # v1_parity.py:15
BOUNDARY_CLOSED = 0          # v1_parity.py:16-20
BOUNDARY_TRANSMISSIVE = 1
BOUNDARY_REFLECTIVE = 2
BOUNDARY_STAGE = 3
BOUNDARY_OUTFLOW = 4

LEDGER_FIELDS = (            # parity_runner.py:38-51 (MassLedger), v1_parity.py:90-103
    "rainfall_input", "infiltration_loss", "boundary_net_flux", "boundary_inflow",
    "boundary_outflow", "external_boundary_inflow", "surface_capture",
    "surface_surcharge", "outfall_discharge", "overflow_surcharge",
    "unreturned_overflow", "recharge_capture", "recharge_loss",
)


class SurfaceTopology(NamedTuple):          # v1_parity.py:23-34
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


class CouplingTopology(NamedTuple):         # v1_parity.py:37-49
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


class PipeTopology(NamedTuple):             # v1_parity.py:52-69
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


class RechargeTopology(NamedTuple):         # v1_parity.py:72-77
    source_cell: np.ndarray
    storage_capacity: np.ndarray
    capture_capacity: np.ndarray
    recharge_capacity: np.ndarray
    stop_depth: np.ndarray


class ParityState(NamedTuple):              # v1_parity.py:80-87
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    wet_time_s: np.ndarray
    node_volume: np.ndarray
    link_flow: np.ndarray
    recharge_storage: np.ndarray


class ParityForcing(NamedTuple):            # parity_inputs.py:36-46
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


def _f(dtype):
    return np.dtype(dtype).type


def cast_tuple(t, dtype):
    """Return ``t`` with every floating array converted to ``dtype``."""
    out = []
    for v in t:
        if isinstance(v, np.ndarray) and np.issubdtype(v.dtype, np.floating):
            out.append(v.astype(dtype))
        else:
            out.append(v)
    return type(t)(*out)


def scatter_add(indices, values, size):
    """v1_parity.py:106-107 -- zeros(size).at[indices].add(values)."""
    out = np.zeros((size,), dtype=values.dtype)
    np.add.at(out, indices, values)
    return out


def _hll(h_l, hu_l, hv_l, h_r, hu_r, hv_r, nx, ny, dtype):
    """v1_parity.py:110-132 -- HLL normal flux; zero where both sides are dry."""
    F = _f(dtype)
    g, dry = F(GRAVITY), F(MIN_DEPTH)
    h_l = np.maximum(h_l, F(0))
    h_r = np.maximum(h_r, F(0))
    wet_l, wet_r = h_l > dry, h_r > dry
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        u_l = np.where(wet_l, hu_l / np.maximum(h_l, dry), F(0))
        v_l = np.where(wet_l, hv_l / np.maximum(h_l, dry), F(0))
        u_r = np.where(wet_r, hu_r / np.maximum(h_r, dry), F(0))
        v_r = np.where(wet_r, hv_r / np.maximum(h_r, dry), F(0))
        un_l, un_r = u_l * nx + v_l * ny, u_r * nx + v_r * ny
        c_l, c_r = np.sqrt(g * h_l), np.sqrt(g * h_r)
        half_g = F(0.5 * GRAVITY)
        p_l, p_r = half_g * (h_l * h_l), half_g * (h_r * h_r)
        f_l = np.stack((h_l * un_l, hu_l * un_l + p_l * nx, hv_l * un_l + p_l * ny))
        f_r = np.stack((h_r * un_r, hu_r * un_r + p_r * nx, hv_r * un_r + p_r * ny))
        q_l, q_r = np.stack((h_l, hu_l, hv_l)), np.stack((h_r, hu_r, hv_r))
        s_l = np.minimum(un_l - c_l, un_r - c_r)
        s_r = np.maximum(un_l + c_l, un_r + c_r)
        denom = np.where(np.abs(s_r - s_l) > F(1.0e-12), s_r - s_l, F(1))
        middle = (s_r * f_l - s_l * f_r + s_l * s_r * (q_r - q_l)) / denom
        flux = np.where(s_l[None] >= F(0), f_l, np.where(s_r[None] <= F(0), f_r, middle))
    return np.where((wet_l | wet_r)[None], flux, F(0)).astype(dtype)


def _hydrostatic_flux_pair(h_l, hu_l, hv_l, z_l, h_r, hu_r, hv_r, z_r, nx, ny, dtype):
    """v1_parity.py:135-150 -- Audusse hydrostatic reconstruction + HLL."""
    F = _f(dtype)
    eps = F(MIN_DEPTH)
    h_l, h_r = np.maximum(h_l, F(0)), np.maximum(h_r, F(0))
    z_star = np.maximum(z_l, z_r)
    hs_l = np.maximum(h_l + z_l - z_star, F(0))
    hs_r = np.maximum(h_r + z_r - z_star, F(0))
    wet_l, wet_r = h_l > eps, h_r > eps
    with np.errstate(divide="ignore", invalid="ignore"):
        u_l = np.where(wet_l, hu_l / np.maximum(h_l, eps), F(0))
        v_l = np.where(wet_l, hv_l / np.maximum(h_l, eps), F(0))
        u_r = np.where(wet_r, hu_r / np.maximum(h_r, eps), F(0))
        v_r = np.where(wet_r, hv_r / np.maximum(h_r, eps), F(0))
    flux = _hll(hs_l, u_l * hs_l, v_l * hs_l, hs_r, u_r * hs_r, v_r * hs_r, nx, ny, dtype)
    half_g = F(0.5 * GRAVITY)
    pc_l = half_g * (h_l * h_l - hs_l * hs_l)
    pc_r = half_g * (h_r * h_r - hs_r * hs_r)
    zero = np.zeros_like(pc_l)
    corr_l = np.stack((zero, pc_l * nx, pc_l * ny))
    corr_r = np.stack((zero, pc_r * nx, pc_r * ny))
    return flux + corr_l, flux + corr_r


def stable_timestep(state: ParityState, topology: SurfaceTopology, cfl=0.85, dtype=np.float32):
    """v1_parity.py:153-158 -- min over all cells of cfl * Lc / (|u| + sqrt(g h))."""
    F = _f(dtype)
    eps = F(MIN_DEPTH)
    h_safe = np.maximum(state.h, eps)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        u = state.hu / h_safe
        v = state.hv / h_safe
        speed = np.sqrt(u * u + v * v)
        speed = speed + np.sqrt(F(GRAVITY) * h_safe)
        local = F(cfl) * topology.characteristic_length / np.maximum(speed, F(1.0e-12))
    local = np.where(np.isfinite(local), local, F(np.inf))
    return F(np.min(local)) if local.size else F(np.inf)


def surface_flux_step(state: ParityState, topology: SurfaceTopology, dt, dtype=np.float32,
                      boundary_stage_m=None):
    """v1_parity.py:161-214 -- edge fluxes, scatter, cell update, Manning friction.

    ``boundary_stage_m`` (a scalar or one value per edge) turns code-3 edges into
    the V1 Dirichlet stage ghost of triangular_backend.py:1294-1327 (CUDA form
    triangular_swe_cuda_kernels.py:86-93): depth max(stage - z_left, 0), zero
    momentum, the left cell's bed.  None keeps v1_parity.py's ghost, which copies
    the left state at code 3 (v1_parity.py:182-186 has no stage values).

    Returns (state, boundary_net, boundary_in, boundary_out).
    """
    F = _f(dtype)
    dt = F(dt)
    g, eps = F(GRAVITY), F(MIN_DEPTH)
    left, right = topology.edge_left, topology.edge_right
    active = topology.active > 0.5
    interior = right >= 0
    safe_right = np.where(interior, right, left)
    left_active = active[left]
    right_active = active[safe_right]
    active_interior = interior & left_active & right_active
    inactive_right = interior & left_active & ~right_active
    code = np.where(inactive_right, BOUNDARY_CLOSED, topology.boundary_code)
    zero = F(0)
    h_l = np.where(left_active, state.h[left], zero)
    hu_l = np.where(left_active, state.hu[left], zero)
    hv_l = np.where(left_active, state.hv[left], zero)
    z_l = topology.bed[left]
    h_rb = np.where(right_active, state.h[safe_right], zero)
    hu_rb = np.where(right_active, state.hu[safe_right], zero)
    hv_rb = np.where(right_active, state.hv[safe_right], zero)
    z_rb = topology.bed[safe_right]
    nx, ny = topology.edge_nx, topology.edge_ny
    normal_momentum = hu_l * nx + hv_l * ny
    hu_reflect = hu_l - F(2) * normal_momentum * nx
    hv_reflect = hv_l - F(2) * normal_momentum * ny
    reflective = (code == BOUNDARY_CLOSED) | (code == BOUNDARY_REFLECTIVE)
    h_ghost, hu_ghost, hv_ghost = h_l, np.where(reflective, hu_reflect, hu_l), np.where(reflective, hv_reflect, hv_l)
    if boundary_stage_m is not None:
        stage = np.broadcast_to(np.asarray(boundary_stage_m, dtype=dtype), code.shape)
        at_stage = code == BOUNDARY_STAGE
        h_ghost = np.where(at_stage, np.maximum(stage - z_l, zero), h_ghost)
        hu_ghost = np.where(at_stage, zero, hu_ghost)
        hv_ghost = np.where(at_stage, zero, hv_ghost)
    h_r = np.where(active_interior, h_rb, h_ghost)
    hu_r = np.where(active_interior, hu_rb, hu_ghost)
    hv_r = np.where(active_interior, hv_rb, hv_ghost)
    z_r = np.where(active_interior, z_rb, z_l)
    flux_l, flux_r = _hydrostatic_flux_pair(
        h_l, hu_l, hv_l, z_l, h_r, hu_r, hv_r, z_r, nx, ny, dtype)
    outflow_edge = (~interior) & (code == BOUNDARY_OUTFLOW)
    blocked_inflow = outflow_edge & (flux_l[0] < zero)
    flux_l = np.where(blocked_inflow[None], zero, flux_l)
    edge_length = topology.edge_length
    delta_l = -flux_l * edge_length
    delta_r = np.where(active_interior[None], flux_r * edge_length, zero)
    count = topology.cell_area.shape[0]
    safe_scatter = np.where(interior, right, 0)
    delta = np.stack([
        scatter_add(left, delta_l[i], count) + scatter_add(safe_scatter, delta_r[i], count)
        for i in range(3)
    ])
    area = topology.cell_area
    h = np.where(active, np.maximum(state.h + dt * delta[0] / area, zero), zero)
    hu = np.where(active, state.hu + dt * delta[1] / area, zero)
    hv = np.where(active, state.hv + dt * delta[2] / area, zero)
    wet = h > eps
    hs = np.maximum(h, eps)
    with np.errstate(over="ignore", invalid="ignore"):
        u = hu / hs
        v = hv / hs
        speed = np.sqrt(u * u + v * v)
        n2 = topology.manning_n * topology.manning_n
        damping = F(1) + dt * g * n2 * speed / hs ** F(4.0 / 3.0)
    hu, hv = np.where(wet, hu / damping, zero), np.where(wet, hv / damping, zero)
    boundary = ~active_interior
    outward = np.where(boundary, flux_l[0] * edge_length * dt, zero).astype(np.float64)
    boundary_in = float(np.sum(np.maximum(-outward, 0.0)))
    boundary_out = float(np.sum(np.maximum(outward, 0.0)))
    return (state._replace(h=h.astype(dtype), hu=hu.astype(dtype), hv=hv.astype(dtype)),
            boundary_in - boundary_out, boundary_in, boundary_out)


def apply_rainfall_horton(state, topology, rainfall_rate_mps, rainfall_multiplier,
                          horton_f0_mps, horton_fc_mps, horton_k_per_s, dt, dtype=np.float32):
    """v1_parity.py:217-242 -- rain, Horton infiltration, momentum rescale."""
    F = _f(dtype)
    dt = F(dt)
    eps = F(MIN_DEPTH)
    active = topology.active.astype(dtype)
    rain_rate = np.maximum(F(rainfall_rate_mps), F(0)) * rainfall_multiplier
    rain_depth = rain_rate * dt * active
    h_before = state.h + rain_depth
    capacity = horton_fc_mps + (horton_f0_mps - horton_fc_mps) * np.exp(-horton_k_per_s * state.wet_time_s)
    infiltration = np.minimum(h_before, np.maximum(capacity, F(0)) * dt * active)
    h = h_before - infiltration
    scale = np.where(h_before > eps, h / np.maximum(h_before, eps), F(0))
    hu = np.where(h > eps, state.hu * scale, F(0))
    hv = np.where(h > eps, state.hv * scale, F(0))
    area64 = topology.cell_area.astype(np.float64)
    return (state._replace(h=h.astype(dtype), hu=hu.astype(dtype), hv=hv.astype(dtype)),
            float(np.sum(rain_depth.astype(np.float64) * area64)),
            float(np.sum(infiltration.astype(np.float64) * area64)))


def clamp_velocity(state, maximum_speed=5.0, dtype=np.float32):
    """v1_parity.py:245-256 -- cap |u| at maximum_speed."""
    F = _f(dtype)
    eps = F(MIN_DEPTH)
    wet = state.h > eps
    with np.errstate(over="ignore", invalid="ignore"):
        speed = np.where(wet, np.sqrt(state.hu * state.hu + state.hv * state.hv) / np.maximum(state.h, eps), F(0))
        scale = np.minimum(F(1), F(maximum_speed) / np.maximum(speed, F(1.0e-12)))
    return state._replace(hu=np.where(wet, state.hu * scale, F(0)).astype(dtype),
                          hv=np.where(wet, state.hv * scale, F(0)).astype(dtype))


def _limit_by_group(requested, group, available, group_count, dtype):
    """v1_parity.py:259-263 -- donor limiting: scale requests so sum <= available."""
    F = _f(dtype)
    requested = np.maximum(requested, F(0))
    total = scatter_add(group, requested, group_count)
    factor = np.minimum(F(1), available / np.maximum(total, F(1.0e-12)))
    return requested * factor[group]


def _node_head(volume, coupling, dtype):
    """v1_parity.py:266-267."""
    F = _f(dtype)
    return coupling.node_invert + np.maximum(volume, F(0)) / np.maximum(coupling.node_storage_area, F(1.0e-12))


def inlet_exchange(state, surface, coupling, dt, dtype=np.float32,
                   capture_q=0.62 * 0.04):
    """v1_parity.py:270-302 -- orifice exchange between surface cells and nodes."""
    F = _f(dtype)
    dt = F(dt)
    cell, node = coupling.inlet_cell, coupling.inlet_node
    node_head = _node_head(state.node_volume, coupling, dtype)
    surface_head = surface.bed[cell] + state.h[cell]
    delta_head = surface_head - node_head[node]
    exchange_q = F(capture_q) * np.sqrt(F(2.0 * GRAVITY) * np.abs(delta_head)) * np.sign(delta_head)
    requested_capture = np.minimum(np.maximum(exchange_q, F(0)) * dt,
                                   np.maximum(coupling.inlet_capture_capacity, F(0)) * dt)
    requested_surcharge = np.minimum(np.maximum(-exchange_q, F(0)) * dt,
                                     np.maximum(coupling.inlet_surcharge_capacity, F(0)) * dt)
    available_by_cell = state.h * surface.cell_area
    capture = _limit_by_group(requested_capture, cell, available_by_cell, surface.cell_area.shape[0], dtype)
    surcharge = _limit_by_group(requested_surcharge, node, state.node_volume,
                                coupling.node_storage_area.shape[0], dtype)
    cell_exchange = scatter_add(cell, (surcharge - capture) / surface.cell_area[cell], surface.cell_area.shape[0])
    node_exchange = scatter_add(node, capture - surcharge, coupling.node_storage_area.shape[0])
    return (state._replace(h=np.maximum(state.h + cell_exchange, F(0)).astype(dtype),
                           node_volume=np.maximum(state.node_volume + node_exchange, F(0)).astype(dtype)),
            float(np.sum(capture.astype(np.float64))), float(np.sum(surcharge.astype(np.float64))))


def _circular_theta(depth, diameter, dtype):
    F = _f(dtype)
    radius = F(0.5) * diameter
    open_depth = np.minimum(np.maximum(depth, F(0)), diameter)
    return radius, F(2) * np.arccos(np.clip((radius - open_depth) / radius, F(-1), F(1)))


def _circular_area(depth, diameter, dtype):
    """v1_parity.py:305-312 -- circular segment area, full area once surcharged."""
    F = _f(dtype)
    diameter = np.maximum(diameter, F(1.0e-12))
    radius, theta = _circular_theta(depth, diameter, dtype)
    open_area = F(0.5) * (radius * radius) * (theta - np.sin(theta))
    full = F(0.25 * np.pi) * (diameter * diameter)
    return np.where(depth >= diameter, full, open_area)


def _circular_radius(depth, diameter, dtype):
    """v1_parity.py:315-321 -- hydraulic radius A / P."""
    F = _f(dtype)
    diameter = np.maximum(diameter, F(1.0e-12))
    radius, theta = _circular_theta(depth, diameter, dtype)
    perimeter = np.where(depth >= diameter, F(np.pi) * diameter, radius * theta)
    return _circular_area(depth, diameter, dtype) / np.maximum(perimeter, F(1.0e-12))


def pipe_links_and_outfalls(state, coupling, pipe, dt, dtype=np.float32, slot_wave_speed=30.0,
                            with_transfer=False):
    """v1_parity.py:324-368 -- semi-implicit link momentum, donor-limited transfer, outfalls.
    ``with_transfer`` also returns the float64 sum of |transfer| over the links."""
    F = _f(dtype)
    dt = F(dt)
    g = F(GRAVITY)
    n_nodes = state.node_volume.shape[0]
    transfer_abs = 0.0
    if pipe.link_from.size == 0 and n_nodes == 0:
        return (state, 0.0, 0.0) if with_transfer else (state, 0.0)
    head = _node_head(state.node_volume, coupling, dtype)
    source, target = pipe.link_from, pipe.link_to
    node_volume = state.node_volume
    link_flow = state.link_flow
    if source.size:
        D = pipe.link_diameter
        depth = np.maximum(F(0.5) * (head[source] + head[target]) - pipe.link_invert, F(0))
        Dm = np.maximum(D, F(1.0e-12))
        slot_width = g * (F(0.25 * np.pi) * (Dm * Dm)) / F(slot_wave_speed ** 2)
        circular_area = _circular_area(depth, D, dtype)
        flow_area = circular_area + slot_width * np.maximum(depth - D, F(0))
        circular_radius = _circular_radius(depth, D, dtype)
        open_depth = np.minimum(np.maximum(depth, F(0)), np.maximum(pipe.link_rect_height, F(1.0e-6)))
        open_area = np.maximum(pipe.link_rect_width, F(1.0e-6)) * open_depth
        open_radius = open_area / np.maximum(pipe.link_rect_width + F(2) * open_depth, F(1.0e-12))
        rect = pipe.link_geometry_code == 2
        flow_area = np.where(rect, open_area, flow_area)
        radius = np.where(rect, open_radius, circular_radius)
        length = np.maximum(pipe.link_length, F(1.0e-6))
        area = np.maximum(flow_area, F(1.0e-12))
        radius = np.maximum(radius, F(1.0e-12))
        slope = (head[source] - head[target]) / length
        nman = np.maximum(pipe.link_manning, F(0))
        absq = np.abs(state.link_flow)
        friction = g * (nman * nman) * absq / (radius ** F(4.0 / 3.0) * area)
        minor = np.maximum(pipe.link_minor_loss, F(0)) * absq / (F(2) * length * area)
        candidate = (state.link_flow + g * area * dt * slope) / (F(1) + dt * (friction + minor))
        cap = np.maximum(pipe.link_capacity, F(0))
        candidate = np.clip(candidate, -cap, cap)
        requested = candidate * dt
        positive = requested >= F(0)
        donor = np.where(positive, source, target)
        requested_abs = np.abs(requested)
        by_donor = scatter_add(donor, requested_abs, n_nodes)
        factor = np.minimum(F(1), node_volume / np.maximum(by_donor, F(1.0e-12)))
        transfer = np.where(positive, F(1), F(-1)) * requested_abs * factor[donor]
        node_delta = scatter_add(np.concatenate((source, target)),
                                 np.concatenate((-transfer, transfer)), n_nodes)
        node_volume = node_volume + node_delta
        link_flow = transfer / np.maximum(dt, F(1.0e-12))
        transfer_abs = float(np.sum(np.abs(transfer).astype(np.float64)))
    node_volume = np.maximum(node_volume, F(0))     # applied to every node, links or not
    head = _node_head(node_volume, coupling, dtype)
    dh = head - pipe.outfall_tailwater
    outfall_q = pipe.outfall_coefficient * pipe.outfall_area * np.sqrt(F(2.0 * GRAVITY) * np.abs(dh)) * np.sign(dh)
    outfall_q = np.minimum(outfall_q, np.maximum(pipe.outfall_capacity, F(0)))
    outfall = np.minimum(node_volume, np.maximum(outfall_q, F(0)) * dt)
    node_volume = np.maximum(node_volume - outfall, F(0))
    out = (state._replace(node_volume=node_volume.astype(dtype), link_flow=link_flow.astype(dtype)),
           float(np.sum(outfall.astype(np.float64))))
    return out + (transfer_abs,) if with_transfer else out


DRY_WEATHER_FIELDS = ("dry_weather_inflow", "link_abs_flow", "outfall_discharge", "overflow_surcharge",
                      "final_node_volume")


def dry_weather_step(state, coupling, pipe, dt, node_inflow_m3ps=None, outfall_tailwater_head_m=None,
                     cap_node_overflow=True, dtype=np.float32):
    """gurugram_flood/kernels/triangular_backend.py:833-952 -- sewer-only dry-weather flow:
    node inflow, then the link momentum / donor-limited transfer / outfall update of
    v1_parity.py:324-368 (the same physics as the reference's dynamic pipe graph), then
    the overflow above storage capacity reported and, by default, capped.  No surface
    exchange.  Returns (state, diagnostics over DRY_WEATHER_FIELDS, float64)."""
    F = _f(dtype)
    dt = F(dt)
    volume = np.maximum(state.node_volume, F(0))
    if node_inflow_m3ps is None:
        inflow = volume * F(0)
    else:
        inflow = np.maximum(np.asarray(node_inflow_m3ps, dtype), F(0))
    inflow_m3 = inflow * dt
    volume = volume + inflow_m3
    if outfall_tailwater_head_m is not None:
        pipe = pipe._replace(outfall_tailwater=np.asarray(outfall_tailwater_head_m, dtype))
    st, outfall, transfer_abs = pipe_links_and_outfalls(state._replace(node_volume=volume), coupling, pipe, dt,
                                                        dtype, with_transfer=True)
    volume = st.node_volume
    capacity = coupling.node_storage_area * np.maximum(coupling.node_max_depth, F(0))
    overflow = np.maximum(volume - capacity, F(0))
    if cap_node_overflow:
        volume = np.minimum(volume, capacity)
    diag = dict(zip(DRY_WEATHER_FIELDS, (float(np.sum(inflow_m3.astype(np.float64))), transfer_abs, outfall,
                                         float(np.sum(overflow.astype(np.float64))),
                                         float(np.sum(volume.astype(np.float64))))))
    return st._replace(node_volume=volume.astype(dtype)), diag


def return_node_overflow(state, surface, coupling, dtype=np.float32):
    """v1_parity.py:371-404 -- weighted entry-table return, then node-surface-cell fallback."""
    F = _f(dtype)
    n_nodes = state.node_volume.shape[0]
    n_cells = surface.cell_area.shape[0]
    if n_nodes == 0:
        return state, 0.0, 0.0
    capacity = coupling.node_storage_area * np.maximum(coupling.node_max_depth, F(0))
    overflow = np.maximum(state.node_volume - capacity, F(0))
    snode, scell = coupling.surcharge_node, coupling.surcharge_cell
    valid = (snode >= 0) & (scell >= 0)
    safe_node = np.where(valid, snode, 0)
    safe_cell = np.where(valid, scell, 0)
    raw_weight = np.where(valid, np.maximum(coupling.surcharge_weight, F(0)), F(0)).astype(dtype)
    weight_sum = scatter_add(safe_node, raw_weight, n_nodes)
    if bool(coupling.surcharge_weight_normalized):
        normalized = raw_weight
    else:
        normalized = np.where(valid, raw_weight / np.maximum(weight_sum[safe_node], F(1.0e-12)), F(0)).astype(dtype)
    returned_entry = overflow[safe_node] * normalized
    h = state.h + scatter_add(safe_cell, returned_entry / surface.cell_area[safe_cell], n_cells)
    returned_node = scatter_add(safe_node, returned_entry, n_nodes)
    returned_node = np.minimum(returned_node, overflow)
    node_volume = state.node_volume - returned_node
    residual = np.maximum(node_volume - capacity, F(0))
    explicit = scatter_add(safe_node, np.where(valid, F(1), F(0)).astype(dtype), n_nodes) > 0
    fallback_valid = (~explicit) & (coupling.node_surface_cell >= 0)
    fallback_cell = np.where(fallback_valid, coupling.node_surface_cell, 0)
    fallback = np.where(fallback_valid, residual, F(0))
    h = h + scatter_add(fallback_cell, fallback / surface.cell_area[fallback_cell], n_cells)
    node_volume = node_volume - fallback
    total = returned_node + fallback
    unreturned = float(np.sum(np.maximum(node_volume - capacity, F(0)).astype(np.float64)))
    return (state._replace(h=h.astype(dtype), node_volume=node_volume.astype(dtype)),
            float(np.sum(total.astype(np.float64))), unreturned)


def apply_external_inflow(state, receiver_cell, q_m3ps, surface, dt, dtype=np.float32):
    """v1_parity.py:407-412 -- add max(q,0) dt / A to receiver cells."""
    F = _f(dtype)
    if receiver_cell.size == 0:
        return state, 0.0
    added = np.maximum(np.asarray(q_m3ps, dtype=dtype), F(0)) * F(dt)
    depth = scatter_add(receiver_cell, added / surface.cell_area[receiver_cell], surface.cell_area.shape[0])
    return state._replace(h=(state.h + depth).astype(dtype)), float(np.sum(added.astype(np.float64)))


def apply_recharge(state, surface, recharge, dt, dtype=np.float32):
    """v1_parity.py:415-427 -- pond/RWH capture into storage, then recharge loss."""
    F = _f(dtype)
    dt = F(dt)
    cell = recharge.source_cell
    if cell.size == 0:
        return state, 0.0, 0.0
    area = np.maximum(surface.cell_area[cell], F(1.0e-6))
    available = np.maximum(state.h[cell] - recharge.stop_depth, F(0)) * area
    capture = np.minimum(np.minimum(available, recharge.capture_capacity * dt),
                         np.maximum(recharge.storage_capacity - state.recharge_storage, F(0)))
    removed_by_cell = scatter_add(cell, capture, surface.cell_area.shape[0])
    h = np.maximum(state.h - removed_by_cell / surface.cell_area, F(0))
    storage = state.recharge_storage + capture
    loss = np.minimum(storage, recharge.recharge_capacity * dt)
    return (state._replace(h=h.astype(dtype), recharge_storage=(storage - loss).astype(dtype)),
            float(np.sum(capture.astype(np.float64))), float(np.sum(loss.astype(np.float64))))


def advance_parity_step(state, surface, coupling, pipe, recharge, *, rainfall_rate_mps,
                        rainfall_multiplier, horton_f0_mps, horton_fc_mps, horton_k_per_s,
                        external_receiver_cell, external_q_m3ps, dt, dtype=np.float32,
                        max_speed=5.0, surface_mode="full_swe", boundary_stage_m=None):
    """v1_parity.py:430-469 -- one step in the V1 process order.

    ``surface_mode="local_inertial"`` swaps the surface operator for the
    reference's local-inertial step (engine.py:253-256, in oracle/local_inertial_np.py);
    the sources, clamp, drainage and ledger are the same.  ``boundary_stage_m``
    gives code-3 edges their stage ghost (surface_flux_step; the local-inertial
    step treats code 3 as a wall either way).
    Returns (state, diagnostics) with diagnostics a dict over LEDGER_FIELDS (float64).
    """
    F = _f(dtype)
    dt = F(dt)
    if surface_mode == "local_inertial":
        from oracle import local_inertial_np
        state, bnet, bin_, bout = local_inertial_np.surface_step(state, surface, dt, dtype)
    elif surface_mode == "full_swe":
        state, bnet, bin_, bout = surface_flux_step(state, surface, dt, dtype, boundary_stage_m)
    else:
        raise ValueError(f"unknown surface_mode {surface_mode!r}")
    state, rain, infiltration = apply_rainfall_horton(
        state, surface, rainfall_rate_mps, rainfall_multiplier,
        horton_f0_mps, horton_fc_mps, horton_k_per_s, dt, dtype)
    state = clamp_velocity(state, max_speed, dtype)
    state, external = apply_external_inflow(state, external_receiver_cell, external_q_m3ps, surface, dt, dtype)
    state, capture, surcharge = inlet_exchange(state, surface, coupling, dt, dtype)
    state, outfall = pipe_links_and_outfalls(state, coupling, pipe, dt, dtype)
    state, overflow, unreturned = return_node_overflow(state, surface, coupling, dtype)
    state, rcap, rloss = apply_recharge(state, surface, recharge, dt, dtype)
    state = state._replace(wet_time_s=(state.wet_time_s + dt * (state.h > F(MIN_DEPTH))).astype(dtype))
    diag = dict(zip(LEDGER_FIELDS, (rain, infiltration, bnet, bin_, bout, external, capture,
                                    surcharge, outfall, overflow, unreturned, rcap, rloss)))
    return state, diag


# ---------------------------------------------------------------------------
# Driver: parity_runner.py:96-158 and :203-234
# ---------------------------------------------------------------------------

def rain_rate(forcing: ParityForcing, time_s):
    """parity_runner.py:96-102 -- piecewise-constant [start, end) intervals."""
    ends = forcing.rainfall_end_s
    n = forcing.rainfall_rate_mps.shape[0]
    if n == 0:
        return 0.0
    index = int(np.searchsorted(ends, np.float64(time_s), side="right"))
    safe = min(max(index, 0), n - 1)
    valid = index < n and time_s >= float(forcing.rainfall_start_s[safe])
    return float(forcing.rainfall_rate_mps[safe]) if valid else 0.0


def external_q(forcing: ParityForcing, time_s):
    """parity_runner.py:105-115 -- linear interpolation per receiver, clamped at the ends."""
    times = forcing.external_times_s
    if times.shape[0] == 0 or forcing.external_receiver_cell.shape[0] == 0:
        return np.zeros((0,), np.float32)
    upper = int(np.searchsorted(times, np.float64(time_s), side="right"))
    lower = min(max(upper - 1, 0), times.shape[0] - 1)
    upper_safe = min(max(upper, 0), times.shape[0] - 1)
    denominator = max(float(times[upper_safe]) - float(times[lower]), 1.0e-12)
    fraction = min(max((time_s - float(times[lower])) / denominator, 0.0), 1.0)
    q = forcing.external_q_m3ps
    return (q[lower] + np.float32(fraction) * (q[upper_safe] - q[lower])).astype(np.float32)


class OracleClock(NamedTuple):
    time_s: float
    cached_dt_s: float
    global_step: int


def choose_dt(state, surface, clock: OracleClock, final_time_s, *, cfl=0.85, interval=32,
              reuse=0.8, max_dt=1.0, dt_floor=1.0e-3, fixed_dt=None, dtype=np.float32,
              surface_mode="full_swe"):
    """parity_runner.py:124-135 -- CFL every `interval` steps, else reuse x0.8; caps; floor.
    The local-inertial mode evaluates the gravity-wave timestep (engine.py:220-225).

    Returns (dt, evaluated_dt).  dt is rounded to ``dtype`` (the state precision).
    """
    F = _f(dtype)
    if fixed_dt is not None:
        return F(fixed_dt), clock.cached_dt_s
    evaluate = clock.global_step % interval == 0
    if surface_mode == "local_inertial":
        from oracle import local_inertial_np
        timestep = local_inertial_np.stable_timestep
    else:
        timestep = stable_timestep
    evaluated = float(timestep(state, surface, cfl, dtype)) if evaluate else clock.cached_dt_s
    candidate = evaluated if evaluate else float(F(F(evaluated) * F(reuse)))
    dt = max(min(min(candidate, max_dt), final_time_s - clock.time_s), dt_floor)
    return F(dt), evaluated


def advance_to_time(state, clock: OracleClock, ledger, target_time_s, final_time_s,
                    surface, coupling, pipe, recharge, forcing, *, max_steps=20000,
                    fixed_dt=None, dtype=np.float32, surface_mode="full_swe", boundary_stage_m=None,
                    **dt_kw):
    """parity_runner.py:203-234 -- step while t < target (and the step budget lasts).

    The clock and ledger are kept in float64 (the FP64 clock/ledger the MI355X
    backend keeps on the device), in place of the reference's float32 Kahan pair.
    """
    ledger = dict(ledger)
    steps = 0
    while clock.time_s < target_time_s and steps < max_steps:
        dt, evaluated = choose_dt(state, surface, clock, final_time_s, fixed_dt=fixed_dt,
                                  dtype=dtype, surface_mode=surface_mode, **dt_kw)
        rate = rain_rate(forcing, clock.time_s)
        q = external_q(forcing, clock.time_s + 0.5 * float(dt))
        state, diag = advance_parity_step(
            state, surface, coupling, pipe, recharge,
            rainfall_rate_mps=rate, rainfall_multiplier=forcing.rainfall_multiplier,
            horton_f0_mps=forcing.horton_f0_mps, horton_fc_mps=forcing.horton_fc_mps,
            horton_k_per_s=forcing.horton_k_per_s,
            external_receiver_cell=forcing.external_receiver_cell, external_q_m3ps=q,
            dt=dt, dtype=dtype, surface_mode=surface_mode, boundary_stage_m=boundary_stage_m)
        for k, v in diag.items():
            ledger[k] = ledger.get(k, 0.0) + v
        clock = OracleClock(clock.time_s + float(dt), evaluated, clock.global_step + 1)
        steps += 1
    return state, clock, ledger


def zero_ledger():
    return {k: 0.0 for k in LEDGER_FIELDS}
