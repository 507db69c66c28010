"""Host-side topology preparation for the MI355X backend (one-off, not per step).

Builds the reference's edge topology from node/triangle arrays with vectorized
NumPy (the reference uses a per-edge Python dict loop,
``gurugram_flood/kernels/triangular_backend.py:244-326, 1815-1869``).  The
output arrays have exactly the reference's meaning and order:

* triangles re-oriented counter-clockwise, cell area from the signed area;
* one record per undirected edge, numbered in order of first appearance when
  cells are visited in index order and each cell's sides in (a,b),(b,c),(c,a)
  order; the first cell to meet an edge is its left cell, the normal is that
  cell's edge vector rotated clockwise (outward from the left cell);
* boundary edges get ``right = -1`` and the boundary mode code, interior -1;
* with an active mask, an interior edge whose left cell is inactive and right
  cell active is swapped (normal flipped);
* characteristic length ``Lc = 2 A / perimeter``.
"""

from __future__ import annotations

import numpy as np

from .types import SurfaceTopology

BOUNDARY_CODES = {"closed": 0, "transmissive": 1, "reflective": 2, "stage": 3, "outflow": 4}


def orient_triangles_ccw(nodes: np.ndarray, triangles: np.ndarray):
    """Counter-clockwise vertex order and cell area (the reference's
    _orient_triangles_ccw, triangular_backend.py:1815-1829): a triangle whose
    signed area is negative has its last two vertices exchanged."""
    tri = np.asarray(triangles, dtype=np.int64)
    corner = nodes[tri]                                   # (C, 3, 2)
    e1, e2 = corner[:, 1] - corner[:, 0], corner[:, 2] - corner[:, 0]
    twice = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    clockwise = twice < 0.0
    oriented = np.where(clockwise[:, None], tri[:, (0, 2, 1)], tri)
    return oriented, 0.5 * np.where(clockwise, -twice, twice)


def build_edges(nodes: np.ndarray, tri: np.ndarray):
    """Vectorized equivalent of triangular_backend.py:1832-1869 (_build_edges)."""
    n_cells = tri.shape[0]
    start = np.stack((tri[:, 0], tri[:, 1], tri[:, 2]), axis=1).reshape(-1)
    end = np.stack((tri[:, 1], tri[:, 2], tri[:, 0]), axis=1).reshape(-1)
    vec = nodes[end] - nodes[start]
    length = np.sqrt(vec[:, 0] * vec[:, 0] + vec[:, 1] * vec[:, 1])
    if np.any(length <= 1.0e-12):
        raise ValueError("triangles contain a zero-length edge.")
    perimeter = length.reshape(n_cells, 3).sum(axis=1)
    lo = np.minimum(start, end).astype(np.int64)
    hi = np.maximum(start, end).astype(np.int64)
    key = lo * (int(nodes.shape[0]) + 1) + hi
    order = np.argsort(key, kind="stable")          # stable: first appearance first
    ks = key[order]
    new_group = np.ones(ks.shape[0], dtype=bool)
    new_group[1:] = ks[1:] != ks[:-1]
    group_id = np.cumsum(new_group) - 1
    group_start = np.flatnonzero(new_group)
    counts = np.diff(np.append(group_start, ks.shape[0]))
    if np.any(counts > 2):
        raise ValueError("triangular topology contains a non-manifold edge shared by more than two cells.")
    first_pos = order[group_start]                   # position (cell*3+side) of first appearance
    second_pos = np.full(group_start.shape, -1, dtype=np.int64)
    has2 = counts == 2
    second_pos[has2] = order[group_start[has2] + 1]
    # number edges in order of first appearance
    edge_order = np.argsort(first_pos, kind="stable")
    first_pos = first_pos[edge_order]
    second_pos = second_pos[edge_order]
    left = (first_pos // 3).astype(np.int32)
    right = np.where(second_pos >= 0, second_pos // 3, -1).astype(np.int32)
    v = vec[first_pos]
    L = length[first_pos]
    nx = (v[:, 1] / L).astype(np.float32)
    ny = (-v[:, 0] / L).astype(np.float32)
    del group_id
    return left, right, nx, ny, L.astype(np.float32), perimeter


# (argument, expected shape given (nodes, triangles) counts, what it must be)
_INPUT_SHAPES = (
    ("nodes_xy", lambda nn, nc: (nn, 2), "an (n, 2) array of node coordinates"),
    ("triangles", lambda nn, nc: (nc, 3), "an (n, 3) array of node indices"),
    ("bed_elevation_m", lambda nn, nc: (nc,), "one value per triangle"),
)


def _per_cell(value, n_cells: int, default: float, what: str) -> np.ndarray:
    """A scalar (or None: ``default``) broadcast to every cell, or one value per cell."""
    v = np.asarray(default if value is None else value, dtype=np.float32)
    if v.ndim == 0:
        return np.full(n_cells, v, dtype=np.float32)
    if v.shape != (n_cells,):
        raise ValueError(f"{what} takes a scalar or one value per triangle, got shape {v.shape}.")
    return v.copy()


def build_surface_topology(nodes_xy, triangles, bed_elevation_m, manning_n=None,
                           active_mask=None, boundary_mode="outflow"):
    """The reference's surface topology from a node/triangle mesh
    (build_triangular_surface_topology, triangular_backend.py:244-326): CCW
    cells, one record per edge (build_edges), the boundary code on every edge
    without a right cell, Manning n (default 0.035), the active mask, and
    Lc = 2 A / perimeter.  An interior edge whose left cell is inactive and right
    cell active is given to the active cell (left and right exchanged, normal
    reversed), so every edge's left cell is active when either is.  Returns
    (SurfaceTopology, oriented triangles)."""
    nodes = np.asarray(nodes_xy, dtype=np.float64)
    tri = np.asarray(triangles, dtype=np.int64)
    bed = np.asarray(bed_elevation_m, dtype=np.float32)
    n_nodes = nodes.shape[0] if nodes.ndim == 2 else -1
    n_cells = tri.shape[0] if tri.ndim == 2 else -1
    for (name, shape_of, meaning), value in zip(_INPUT_SHAPES, (nodes, tri, bed)):
        if value.shape != shape_of(n_nodes, n_cells):
            raise ValueError(f"{name} must be {meaning}; got shape {value.shape}.")
    if tri.size and (tri.min() < 0 or tri.max() >= n_nodes):
        raise ValueError("triangles refer to nodes that nodes_xy does not have.")
    oriented, area = orient_triangles_ccw(nodes, tri)
    if np.any(area <= 1.0e-12):
        raise ValueError("the mesh has degenerate (near-zero-area) triangles.")
    left, right, nx, ny, length, perimeter = build_edges(nodes, oriented)
    code = np.where(right < 0, BOUNDARY_CODES[boundary_mode] if isinstance(boundary_mode, str)
                    else int(boundary_mode), -1).astype(np.int32)
    manning = _per_cell(manning_n, n_cells, 0.035, "manning_n")
    active = np.ones(n_cells, dtype=bool) if active_mask is None else _per_cell(active_mask, n_cells, 1.0,
                                                                                "active_mask") > 0.5
    # hand interior edges of an inactive left cell to their active right cell
    flip = (right >= 0) & ~active[left] & active[np.where(right >= 0, right, 0)]
    left, right = np.where(flip, right, left).astype(np.int32), np.where(flip, left, right).astype(np.int32)
    sign = np.where(flip, np.float32(-1.0), np.float32(1.0))
    nx, ny = (nx * sign).astype(np.float32), (ny * sign).astype(np.float32)
    lc = 2.0 * area / np.maximum(perimeter, 1.0e-12)
    return SurfaceTopology(
        bed=bed, manning_n=manning, cell_area=area.astype(np.float32),
        characteristic_length=lc.astype(np.float32),
        edge_left=left, edge_right=right, edge_nx=nx, edge_ny=ny, edge_length=length,
        boundary_code=code, active=active,
    ), oriented


def cell_centroids(nodes_xy, triangles):
    nodes = np.asarray(nodes_xy, dtype=np.float64)
    tri = np.asarray(triangles, dtype=np.int64)
    return nodes[tri].mean(axis=1)
