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
    """triangular_backend.py:1815-1829."""
    tri = np.array(triangles, dtype=np.int64, copy=True)
    p0, p1, p2 = nodes[tri[:, 0]], nodes[tri[:, 1]], nodes[tri[:, 2]]
    a, b = p1 - p0, p2 - p0
    twice = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    cw = twice < 0.0
    if np.any(cw):
        tmp = tri[cw, 1].copy()
        tri[cw, 1] = tri[cw, 2]
        tri[cw, 2] = tmp
        twice[cw] = -twice[cw]
    return tri, 0.5 * twice


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


def build_surface_topology(nodes_xy, triangles, bed_elevation_m, manning_n=None,
                           active_mask=None, boundary_mode="outflow") -> SurfaceTopology:
    """Mirror of triangular_backend.py:244-326 (build_triangular_surface_topology)."""
    nodes = np.asarray(nodes_xy, dtype=np.float64)
    tri = np.asarray(triangles, dtype=np.int64)
    bed = np.asarray(bed_elevation_m, dtype=np.float32)
    if nodes.ndim != 2 or nodes.shape[1] != 2:
        raise ValueError("nodes_xy must have shape (node_count, 2).")
    if tri.ndim != 2 or tri.shape[1] != 3:
        raise ValueError("triangles must have shape (triangle_count, 3).")
    if bed.shape != (tri.shape[0],):
        raise ValueError("bed_elevation_m must have one value per triangle.")
    if np.any(tri < 0) or np.any(tri >= nodes.shape[0]):
        raise ValueError("triangles contain node indices outside nodes_xy.")
    oriented, area = orient_triangles_ccw(nodes, tri)
    if np.any(area <= 1.0e-12):
        raise ValueError("triangles contain degenerate or near-zero-area cells.")
    left, right, nx, ny, length, perimeter = build_edges(nodes, oriented)
    code = np.full(left.shape, -1, dtype=np.int32)
    code[right < 0] = BOUNDARY_CODES[boundary_mode] if isinstance(boundary_mode, str) else int(boundary_mode)
    if manning_n is None:
        manning = np.full(tri.shape[0], 0.035, dtype=np.float32)
    elif np.isscalar(manning_n):
        manning = np.full(tri.shape[0], float(manning_n), dtype=np.float32)
    else:
        manning = np.asarray(manning_n, dtype=np.float32)
    if active_mask is None:
        active = np.ones(tri.shape[0], dtype=bool)
    else:
        active = np.asarray(active_mask) > 0.5
        interior = right >= 0
        swap = interior & (~active[left]) & active[np.maximum(right, 0)]
        if np.any(swap):
            l0 = left.copy()
            left[swap] = right[swap]
            right[swap] = l0[swap]
            nx[swap] *= -1.0
            ny[swap] *= -1.0
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
