"""CPU oracle for topology preparation (TEST INFRASTRUCTURE ONLY).

Literal restatement of the reference's host-side edge builder, used by tests to
check the product's vectorized builder on small meshes.  Reference (relative to
``HF_CUPY_JAX_COMPLETED_PROJECT/reproduction/source/src/``):

* ``gurugram_flood/kernels/triangular_backend.py:1815-1829`` (_orient_triangles_ccw)
* ``gurugram_flood/kernels/triangular_backend.py:1832-1869`` (_build_edges: one
  record per undirected edge in order of first appearance; the first cell to
  meet the edge is its left cell and the normal is that cell's CCW edge vector
  rotated clockwise, i.e. outward from the left cell)
* ``gurugram_flood/kernels/triangular_backend.py:297-312`` (swap so that an
  active cell is on the left of an active/inactive pair; Lc = 2 A / perimeter)

Pure-Python loops: small meshes only.  Pinned in tests against hand-built
two-triangle and unit-square meshes (closed-form normals, lengths, areas).
"""

from __future__ import annotations

import numpy as np


def orient_ccw(nodes, triangles):
    tri = np.array(triangles, dtype=np.int64, copy=True)
    p0, p1, p2 = nodes[tri[:, 0]], nodes[tri[:, 1]], nodes[tri[:, 2]]
    a, b = p1 - p0, p2 - p0
    twice = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    cw = twice < 0.0
    tri[cw, 1], tri[cw, 2] = tri[cw, 2].copy(), tri[cw, 1].copy()
    twice[cw] = -twice[cw]
    return tri, 0.5 * twice


def build_edges_loop(nodes, tri):
    index = {}
    left, right, nx, ny, length = [], [], [], [], []
    perimeter = np.zeros(tri.shape[0], dtype=np.float64)
    for c, (a, b, d) in enumerate(tri):
        for s, e in ((int(a), int(b)), (int(b), int(d)), (int(d), int(a))):
            v = nodes[e] - nodes[s]
            L = float(np.sqrt(np.dot(v, v)))
            perimeter[c] += L
            key = (s, e) if s < e else (e, s)
            k = index.get(key)
            if k is None:
                index[key] = len(left)
                left.append(c)
                right.append(-1)
                nx.append(float(v[1] / L))
                ny.append(float(-v[0] / L))
                length.append(L)
            else:
                if right[k] >= 0:
                    raise ValueError("non-manifold edge")
                right[k] = c
    return (np.asarray(left, np.int32), np.asarray(right, np.int32),
            np.asarray(nx, np.float32), np.asarray(ny, np.float32),
            np.asarray(length, np.float32), perimeter)


def build_surface_loop(nodes, triangles, bed, manning, active=None, boundary_code_value=4):
    nodes = np.asarray(nodes, np.float64)
    tri, area = orient_ccw(nodes, triangles)
    left, right, nx, ny, length, perim = build_edges_loop(nodes, tri)
    code = np.full(left.shape, -1, np.int32)
    code[right < 0] = boundary_code_value
    if active is not None:
        act = np.asarray(active) > 0.5
        swap = (right >= 0) & (~act[left]) & act[np.maximum(right, 0)]
        l0 = left.copy()
        left[swap] = right[swap]
        right[swap] = l0[swap]
        nx[swap] *= -1.0
        ny[swap] *= -1.0
    lc = 2.0 * area / np.maximum(perim, 1e-12)
    return dict(edge_left=left, edge_right=right, edge_nx=nx, edge_ny=ny, edge_length=length,
                boundary_code=code, cell_area=area.astype(np.float32),
                characteristic_length=lc.astype(np.float32), triangles=tri)
