"""CPU oracle: NumPy restatement of the reference's local-inertial surface step.

TEST INFRASTRUCTURE ONLY.  Nothing on the product path imports this module:
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline
leg use it, and only as the checker.

The reference's CuPy engine runs this surface operator in place of the full
shallow-water step when ``solver.surface_step_mode`` is ``local_inertial``
(``gurugram_flood/solver/engine.py:219-256``; the production configuration's
default, ``inputs/configurations/production_gpu.yaml:11``).  Paths relative to
``HF_CUPY_JAX_COMPLETED_PROJECT/reproduction/source/src/``:

* ``gurugram_flood/kernels/triangular_backend.py:1181-1291``
  (``_advance_triangular_local_inertial``) -- the edge discharge, donor-limited
  interior transfers, boundary outflow capture, continuity and the momentum
  rebuilt from the edge discharges;
* ``gurugram_flood/kernels/local_inertial.py:20-48`` -- the LISFLOOD-FP style
  discharge update;
* ``gurugram_flood/kernels/network_kernels.py:45-82`` -- donor limiting and the
  signed transfer scatter;
* ``gurugram_flood/kernels/triangular_backend.py:1767-1772`` -- capture
  limited by available volume;
* ``gurugram_flood/kernels/equations.py:181-207`` -- the gravity-wave CFL
  timestep.

The vectorized step is the specification; the reference's fused CUDA form of
it (``triangular_local_inertial_cuda_kernels.py:57-300``) documents itself as
an approximation of it.  Where they differ this restatement keeps the
vectorized semantics, in the state precision:

* transfers, their scatter and the new depth stay in the state dtype (the
  vectorized NumPy path promotes them to float64 through
  ``xp.where(positive, 1.0, -1.0)``; the fused CUDA path computes in float32);
* boundary outflow is limited by the volume left after the interior transfers
  (``triangular_backend.py:1257-1264``; the fused path only clips h at 0);
* ``radius ** (4/3)`` is a power (the fused path uses ``r * cbrtf(r)``);
* every per-cell sum runs in ascending edge order (``np.add.at``; the fused
  path's donor sums are float atomics).

Arithmetic follows the reference expression by expression.  Python-float
constants combine with float32 arrays as NumPy does (the constant is rounded
to float32 first), and products of two Python floats (``gravity * dt``) are
formed in double and then rounded.  Scatter-adds use ``np.add.at`` (sequential
in index order).

Pinned by tests/test_oracle_local_inertial.py: lake at rest, closed-basin
volume conservation, a hand-computed two-triangle step, the friction closed
form, donor limiting, the boundary kinds, mirror symmetry, the timestep closed
form and the ledger balance identity of the coupled step.
"""

from __future__ import annotations

import numpy as np

from oracle.v1_parity_np import (BOUNDARY_CLOSED, BOUNDARY_OUTFLOW, BOUNDARY_TRANSMISSIVE, GRAVITY, MIN_DEPTH,
                                 _f, scatter_add)

MIN_FLOW_AREA = 1.0e-8       # local_inertial.py:27
MIN_HYDRAULIC_RADIUS = 1.0e-6  # local_inertial.py:28


def discharge_update(q_old, flow_area, hydraulic_radius, slope, manning_n, dt, dtype=np.float32,
                     gravity=GRAVITY):
    """local_inertial.py:20-48 -- semi-implicit local-inertial discharge."""
    F = _f(dtype)
    area = np.maximum(flow_area, F(MIN_FLOW_AREA))
    radius = np.maximum(hydraulic_radius, F(MIN_HYDRAULIC_RADIUS))
    numerator = q_old - F(gravity) * area * F(dt) * slope
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        denominator = F(1.0) + (F(gravity * float(dt)) * manning_n ** 2 * np.abs(q_old)
                                / (radius ** F(4.0 / 3.0) * area))
        return numerator / denominator


def limit_link_transfers(requested, from_index, to_index, available, dtype=np.float32):
    """network_kernels.py:45-67 -- scale each donor's requests to its available volume."""
    F = _f(dtype)
    positive = requested >= F(0)
    donor = np.where(positive, from_index, to_index)
    requested_abs = np.abs(requested)
    by_donor = scatter_add(donor, requested_abs, available.shape[0])
    factor = np.minimum(F(1), available / np.maximum(by_donor, F(1.0e-12)))
    limited = requested_abs * factor[donor]
    return np.where(positive, F(1), F(-1)) * limited


def delta_from_transfers(transfer, from_index, to_index, count):
    """network_kernels.py:70-82 -- -T at 'from', +T at 'to' (all 'from' entries first)."""
    return scatter_add(np.concatenate([from_index, to_index]), np.concatenate([-transfer, transfer]), count)


def limit_capture(requested, cell_index, available, count, dtype=np.float32):
    """triangular_backend.py:1767-1772 -- per-cell capture limited by available volume."""
    F = _f(dtype)
    requested = np.maximum(requested, F(0))
    by_cell = scatter_add(cell_index, requested, count)
    factor = np.minimum(F(1), available / np.maximum(by_cell, F(1.0e-12)))
    return requested * factor[cell_index]


def stable_timestep(state, topology, cfl=0.85, dtype=np.float32):
    """equations.py:181-207 -- min over all cells of cfl * Lc / sqrt(g max(h, min depth))."""
    F = _f(dtype)
    h_safe = np.maximum(state.h, F(MIN_DEPTH))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        c = np.sqrt(F(GRAVITY) * h_safe)
        local = F(cfl) * topology.characteristic_length / np.maximum(c, F(1.0e-12))
    local = np.where(np.isfinite(local), local, F(np.inf))
    return F(np.min(local)) if local.size else F(np.inf)


def surface_step(state, topology, dt, dtype=np.float32):
    """triangular_backend.py:1181-1291 -- one local-inertial surface step.

    Returns (state, boundary_net, boundary_in, boundary_out) like the full
    shallow-water step of the V1 oracle; the sources, clamp and drainage follow.
    """
    F = _f(dtype)
    dt = F(dt)
    zero = F(0)
    left, right = topology.edge_left, topology.edge_right
    interior = right >= 0
    safe_right = np.where(interior, right, left)
    active = topology.active > 0.5
    left_active = active[left]
    right_active = active[safe_right]
    active_interior = interior & left_active & right_active
    inactive_right = interior & left_active & ~right_active
    code = np.where(inactive_right, BOUNDARY_CLOSED, topology.boundary_code)

    h_l = np.where(left_active, state.h[left], zero)
    hu_l = np.where(left_active, state.hu[left], zero)
    hv_l = np.where(left_active, state.hv[left], zero)
    bed_l = topology.bed[left]
    h_r = np.where(right_active, state.h[safe_right], zero)
    hu_r = np.where(right_active, state.hu[safe_right], zero)
    hv_r = np.where(right_active, state.hv[safe_right], zero)
    bed_r = topology.bed[safe_right]

    eta_l = bed_l + h_l
    eta_r = np.where(active_interior, bed_r + h_r, bed_l)
    bed_star = np.where(active_interior, np.maximum(bed_l, bed_r), bed_l)
    h_flow = np.maximum(np.maximum(eta_l, eta_r) - bed_star, zero)
    length = topology.edge_length
    lc = topology.characteristic_length
    distance = np.maximum(F(0.5) * (lc[left] + lc[safe_right]), F(1.0e-6))
    slope = (eta_r - eta_l) / distance
    nx, ny = topology.edge_nx, topology.edge_ny
    nm_l = hu_l * nx + hv_l * ny
    nm_r = hu_r * nx + hv_r * ny
    nm_face = np.where(active_interior, F(0.5) * (nm_l + nm_r), nm_l)
    q_old = nm_face * length
    n = topology.manning_n
    n_face = np.where(active_interior, F(0.5) * (n[left] + n[safe_right]), n[left])
    can_outflow = (code == BOUNDARY_OUTFLOW) | (code == BOUNDARY_TRANSMISSIVE)
    boundary_edge = (~interior) & left_active & can_outflow
    routable = active_interior | boundary_edge
    q = discharge_update(q_old, h_flow * length, h_flow, slope, n_face, dt, dtype)
    q = np.where(routable, q, zero)
    q = np.where(boundary_edge, np.maximum(q, zero), q)

    count = topology.cell_area.shape[0]
    area = topology.cell_area
    volume = np.maximum(state.h, zero) * area
    requested = np.where(active_interior, q * dt, zero)
    transfer = limit_link_transfers(requested, left, safe_right, volume, dtype)
    interior_delta = delta_from_transfers(transfer, left, safe_right, count)
    volume_after = np.maximum(volume + interior_delta, zero)
    requested_out = np.where(boundary_edge, np.maximum(q, zero) * dt, zero)
    out_by_edge = limit_capture(requested_out, left, volume_after, count, dtype)
    delta = interior_delta - scatter_add(left, out_by_edge, count)
    h = np.where(active, np.maximum(state.h + delta / area, zero), zero)

    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        perimeter = F(2.0) * area / np.maximum(lc, F(1.0e-12))
    qx = q * nx
    qy = q * ny
    hu_sum = scatter_add(left, qx, count) + scatter_add(safe_right, np.where(active_interior, qx, zero), count)
    hv_sum = scatter_add(left, qy, count) + scatter_add(safe_right, np.where(active_interior, qy, zero), count)
    wet = h > F(MIN_DEPTH)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        hu = np.where(wet, hu_sum / np.maximum(perimeter, F(1.0e-12)), zero)
        hv = np.where(wet, hv_sum / np.maximum(perimeter, F(1.0e-12)), zero)
    boundary_out = float(np.sum(out_by_edge.astype(np.float64)))
    boundary_in = boundary_out * 0.0
    return (state._replace(h=h.astype(dtype), hu=hu.astype(dtype), hv=hv.astype(dtype)),
            boundary_in - boundary_out, boundary_in, boundary_out)
