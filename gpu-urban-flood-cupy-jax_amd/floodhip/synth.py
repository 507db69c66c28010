"""Seeded synthetic workloads (the reference's production assets are not available).

Each ``make_*`` returns a :class:`floodhip.types.LoadedParityInputs` with the
reference's array meanings (``hybrid_flood/operational/parity_inputs.py:130-220``):
surface/coupling/pipe/recharge topology, initial state and forcing.  Seeds
derive from ``BASE_SEED + workload index``.

* ``make_plain_small``      -- 100x100 split-square grid over 2 km x 2 km (20,000 triangles),
                               0.5 % slope + Gaussian bump, n = 0.03, 50 mm/h rain for 600 s,
                               no infiltration, no pipes, outflow-only exterior (code 4).
* ``make_production_scale`` -- Delaunay of a jittered 701x701 grid + 940 interior points over
                               17.3 km (981,880 triangles), 6-octave value-noise bed 0-120 m,
                               5 land-cover classes (Manning n, Horton f0/fc/k), pipe network
                               of 118,768 nodes / 139,798 links / 61,317 inlets / 46 outfalls,
                               hourly hyetograph scaled to 160 mm over 24 h (rain in hour 1).
* ``make_dam_break``        -- 10 km x 200 m channel, 200,000 triangles, flat frictionless bed,
                               10 m | 0.001 m dam at x = 5 km, reflective walls.
* ``make_catchment``        -- one member of the uneven ensemble (log-uniform 20k-300k triangles).
* ``make_shared_mesh_ensemble`` -- members on one mesh and network with their own rain total
                               (evenly spaced 40-200 mm over 1 h) and Horton f0/fc/k factors 1+U(-0.2,0.2).
* ``make_uneven_ensemble``  -- 48 catchments with log-uniform 20k-300k triangles, own hyetographs
                               (20-200 mm over 1 h) and networks at ~12 % of their triangle count.
"""

from __future__ import annotations

import numpy as np

from .topology import build_surface_topology, cell_centroids
from .types import (CouplingTopology, LoadedParityInputs, ParityForcing, ParityState,
                    PipeTopology, RechargeTopology, empty_coupling, empty_pipe, empty_recharge)

BASE_SEED = 20261203
MM_HR_TO_M_S = 1.0e-3 / 3600.0


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------

def split_square_grid(n: int, extent: float):
    """n x n squares over [0, extent]^2, each split into two triangles."""
    xs = np.linspace(0.0, extent, n + 1)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    nodes = np.column_stack((X.ravel(), Y.ravel()))
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = (j * (n + 1) + i).ravel()
    b = a + 1
    c = a + (n + 1) + 1
    d = a + (n + 1)
    tri = np.empty((2 * n * n, 3), dtype=np.int64)
    tri[0::2] = np.column_stack((a, b, c))
    tri[1::2] = np.column_stack((a, c, d))
    return nodes, tri


def jittered_delaunay(nx: int, ny: int, width: float, height: float, rng,
                      jitter: float = 0.3, extra_points: int = 0):
    """Delaunay triangulation of an (nx+1) x (ny+1) grid, interior points jittered
    by +-jitter*spacing, boundary points fixed, plus ``extra_points`` seeded interior
    points (each adds two triangles)."""
    from scipy.spatial import Delaunay
    sx, sy = width / nx, height / ny
    gx, gy = np.meshgrid(np.arange(nx + 1) * sx, np.arange(ny + 1) * sy, indexing="xy")
    pts = np.column_stack((gx.ravel(), gy.ravel()))
    ix, iy = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="xy")
    interior = ((ix > 0) & (ix < nx) & (iy > 0) & (iy < ny)).ravel()
    k = int(interior.sum())
    pts[interior, 0] += rng.uniform(-jitter, jitter, k) * sx
    pts[interior, 1] += rng.uniform(-jitter, jitter, k) * sy
    if extra_points:
        ex = np.column_stack((rng.uniform(0.05 * width, 0.95 * width, extra_points),
                              rng.uniform(0.05 * height, 0.95 * height, extra_points)))
        pts = np.vstack((pts, ex))
    tri = Delaunay(pts).simplices.astype(np.int64)
    expected = 2 * pts.shape[0] - 2 * (nx + ny) - 2
    if tri.shape[0] != expected:
        raise RuntimeError(f"Delaunay produced {tri.shape[0]} triangles, expected {expected}")
    return pts, tri


def value_noise(x, y, extent, rng, octaves=6, persistence=0.5, lacunarity=2.0, base_cells=4):
    """Seeded fractal value noise in [0, 1] (smoothstep-interpolated lattice values)."""
    total = np.zeros_like(x, dtype=np.float64)
    amp, freq, norm = 1.0, float(base_cells), 0.0
    for _ in range(octaves):
        f = int(round(freq))
        lattice = rng.random((f + 2, f + 2))
        gx = np.clip(x / extent, 0.0, 1.0) * f
        gy = np.clip(y / extent, 0.0, 1.0) * f
        i0 = np.minimum(np.floor(gx).astype(np.int64), f)
        j0 = np.minimum(np.floor(gy).astype(np.int64), f)
        tx, ty = gx - i0, gy - j0
        tx = tx * tx * (3.0 - 2.0 * tx)
        ty = ty * ty * (3.0 - 2.0 * ty)
        v00, v10 = lattice[j0, i0], lattice[j0, i0 + 1]
        v01, v11 = lattice[j0 + 1, i0], lattice[j0 + 1, i0 + 1]
        total += amp * ((v00 * (1 - tx) + v10 * tx) * (1 - ty) + (v01 * (1 - tx) + v11 * tx) * ty)
        norm += amp
        amp *= persistence
        freq *= lacunarity
    total /= norm
    lo, hi = total.min(), total.max()
    return (total - lo) / max(hi - lo, 1e-12)


# ---------------------------------------------------------------------------
# forcing helpers
# ---------------------------------------------------------------------------

def hourly_hyetograph(total_mm: float, hours: int, rng, peak_hour: float = 3.0):
    """Hourly intensities (mm/h) of a skewed storm, scaled to ``total_mm``, wet from hour 1."""
    t = np.arange(hours) + 0.5
    shape = (t / peak_hour) ** 2 * np.exp(-2.0 * (t / peak_hour - 1.0))
    shape *= rng.uniform(0.8, 1.2, hours)
    shape = np.maximum(shape, 0.02 * shape.max())
    return shape / shape.sum() * total_mm


def rain_forcing(intensity_mm_h, interval_s: float, n_cells: int, multiplier=None,
                 f0=None, fc=None, k=None) -> ParityForcing:
    n = len(intensity_mm_h)
    start = (np.arange(n) * interval_s).astype(np.float32)
    end = (np.arange(1, n + 1) * interval_s).astype(np.float32)
    z = np.zeros(n_cells, np.float32)
    return ParityForcing(
        start, end, (np.asarray(intensity_mm_h, np.float32) * np.float32(MM_HR_TO_M_S)).astype(np.float32),
        np.ones(n_cells, np.float32) if multiplier is None else np.asarray(multiplier, np.float32),
        z if f0 is None else np.asarray(f0, np.float32),
        z.copy() if fc is None else np.asarray(fc, np.float32),
        z.copy() if k is None else np.asarray(k, np.float32),
        np.zeros((0,), np.float32), np.zeros((0,), np.int32), np.zeros((0, 0), np.float32))


# ---------------------------------------------------------------------------
# drainage network
# ---------------------------------------------------------------------------

def pipe_network(rng, centroids, bed, n_nodes: int, n_links: int, n_inlets: int, n_outfalls: int,
                 extent: float, cell_area=None):
    """Seeded drainage graph: spanning tree over nearest neighbours plus loop links."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components, minimum_spanning_tree
    from scipy.spatial import cKDTree

    n_cells = centroids.shape[0]
    node_cell = rng.choice(n_cells, size=n_nodes, replace=False).astype(np.int32)
    pos = centroids[node_cell]
    tree = cKDTree(pos)
    k = min(9, n_nodes)
    dist, nb = tree.query(pos, k=k)
    rows = np.repeat(np.arange(n_nodes), k - 1)
    cols = nb[:, 1:].ravel()
    w = dist[:, 1:].ravel() + 1e-3
    g = coo_matrix((w, (rows, cols)), shape=(n_nodes, n_nodes)).tocsr()
    g = g.maximum(g.T)
    ncomp, labels = connected_components(g, directed=False)
    if ncomp > 1:  # chain components together through their first nodes
        firsts = np.array([np.flatnonzero(labels == c)[0] for c in range(ncomp)])
        extra = coo_matrix((np.full(ncomp - 1, 1e6), (firsts[:-1], firsts[1:])), shape=g.shape).tocsr()
        g = g.maximum(extra).maximum(extra.T)
    mst = minimum_spanning_tree(g).tocoo()
    a, b = mst.row.astype(np.int64), mst.col.astype(np.int64)
    tree_keys = set(zip(np.minimum(a, b).tolist(), np.maximum(a, b).tolist()))
    need = n_links - a.shape[0]
    cand_a, cand_b = rows.astype(np.int64), cols.astype(np.int64)
    lo, hi = np.minimum(cand_a, cand_b), np.maximum(cand_a, cand_b)
    keys = np.unique(lo * n_nodes + hi)
    keys = keys[~np.isin(keys, np.fromiter((x * n_nodes + y for x, y in tree_keys), np.int64, len(tree_keys)))]
    pick = rng.choice(keys.shape[0], size=min(need, keys.shape[0]), replace=False)
    la = np.concatenate((a, keys[pick] // n_nodes))
    lb = np.concatenate((b, keys[pick] % n_nodes))

    depth = rng.uniform(1.5, 4.0, n_nodes).astype(np.float32)
    invert = (bed[node_cell] - depth).astype(np.float32)
    storage_area = np.exp(rng.uniform(np.log(2.0), np.log(30.0), n_nodes)).astype(np.float32)
    max_depth = depth.copy()
    # orient links downhill (from the higher invert to the lower)
    swap = invert[la] < invert[lb]
    link_from = np.where(swap, lb, la).astype(np.int32)
    link_to = np.where(swap, la, lb).astype(np.int32)
    L = n_links = link_from.shape[0]
    length = np.maximum(np.linalg.norm(pos[link_from] - pos[link_to], axis=1), 5.0).astype(np.float32)
    diameter = np.exp(rng.uniform(np.log(0.3), np.log(2.0), L)).astype(np.float32)
    manning = rng.uniform(0.013, 0.016, L).astype(np.float32)
    area = (0.25 * np.pi * diameter ** 2).astype(np.float32)
    radius = (diameter / 4.0).astype(np.float32)
    link_invert = np.minimum(invert[link_from], invert[link_to]).astype(np.float32)
    slope = np.maximum(np.abs(invert[link_from] - invert[link_to]) / length, 1e-3)
    capacity = (0.7 / manning * area * radius ** (2.0 / 3.0) * np.sqrt(slope)).astype(np.float32)
    minor = rng.uniform(0.0, 1.0, L).astype(np.float32)
    geom = np.where(rng.random(L) < 0.1, 2, 1).astype(np.int32)
    rect_w = np.where(geom == 2, diameter, 0.0).astype(np.float32)
    rect_h = np.where(geom == 2, diameter, 0.0).astype(np.float32)

    # outfalls: lowest-invert nodes near the domain boundary
    margin = np.minimum.reduce([pos[:, 0], pos[:, 1], extent - pos[:, 0], extent - pos[:, 1]])
    near = np.flatnonzero(margin < 0.05 * extent)
    if near.shape[0] < n_outfalls:
        near = np.argsort(margin)[: max(n_outfalls, 1)]
    out_nodes = near[np.argsort(invert[near])[:n_outfalls]]
    outfall_area = np.zeros(n_nodes, np.float32)
    outfall_coef = np.zeros(n_nodes, np.float32)
    outfall_tw = np.zeros(n_nodes, np.float32)
    outfall_cap = np.zeros(n_nodes, np.float32)
    outfall_area[out_nodes] = (0.25 * np.pi * rng.uniform(1.0, 2.0, out_nodes.shape[0]) ** 2)
    outfall_coef[out_nodes] = 0.62
    outfall_tw[out_nodes] = invert[out_nodes] + 0.5
    outfall_cap[out_nodes] = 20.0

    # inlets: near a node, snapped to the nearest cell
    inlet_node = rng.integers(0, n_nodes, n_inlets).astype(np.int32)
    ipos = pos[inlet_node] + rng.normal(0.0, 15.0, (n_inlets, 2))
    ipos = np.clip(ipos, 0.0, extent)
    _, inlet_cell = cKDTree(centroids).query(ipos)
    inlet_cell = inlet_cell.astype(np.int32)
    cap_capacity = rng.uniform(0.02, 0.2, n_inlets).astype(np.float32)
    sur_capacity = (cap_capacity * rng.uniform(0.5, 1.0, n_inlets)).astype(np.float32)

    coupling = CouplingTopology(
        inlet_cell, inlet_node, node_cell, invert, storage_area, max_depth,
        cap_capacity, sur_capacity, np.zeros(0, np.int32), np.zeros(0, np.int32),
        np.zeros(0, np.float32), False)
    pipe = PipeTopology(link_from, link_to, length, area, radius, manning, capacity, link_invert,
                        diameter, minor, geom, rect_w, rect_h, outfall_area, outfall_coef,
                        outfall_tw, outfall_cap)
    initial_volume = (storage_area * np.float32(0.5) * max_depth).astype(np.float32)
    return coupling, pipe, initial_volume


def _assemble(surface, coupling, pipe, recharge, forcing, node_volume=None, h0=None, host=None,
              report=None):
    C = surface.cell_area.shape[0]
    nn = coupling.node_storage_area.shape[0]
    nl = pipe.link_from.shape[0]
    nr = recharge.source_cell.shape[0]
    z = np.zeros(C, np.float32)
    state = ParityState(z.copy() if h0 is None else np.asarray(h0, np.float32), z.copy(), z.copy(),
                        z.copy(),
                        np.zeros(nn, np.float32) if node_volume is None else node_volume,
                        np.zeros(nl, np.float32), np.zeros(nr, np.float32))
    return LoadedParityInputs(surface, coupling, pipe, recharge, state, forcing, host or {}, report or {})


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

def make_plain_small(n: int = 100, extent: float = 2000.0, seed: int = BASE_SEED + 1,
                     rain_mm_h: float = 50.0, rain_s: float = 600.0, manning: float = 0.03,
                     active_mask=None):
    """Plain method, small: split-square grid, planar slope + Gaussian bump, uniform rain."""
    rng = np.random.default_rng(seed)
    nodes, tri = split_square_grid(n, extent)
    cen = cell_centroids(nodes, tri)
    p0 = rng.uniform(0.25 * extent, 0.75 * extent, 2)
    sigma = 0.1 * extent
    bed = 0.005 * (extent - cen[:, 0]) + 3.0 * np.exp(-((cen - p0) ** 2).sum(axis=1) / (2 * sigma ** 2))
    surface, tri = build_surface_topology(nodes, tri, bed.astype(np.float32), manning,
                                          active_mask=active_mask, boundary_mode="outflow")
    C = tri.shape[0]
    forcing = rain_forcing([rain_mm_h], rain_s, C)
    host = {"nodes": nodes, "triangles": tri, "centroids": cen}
    return _assemble(surface, empty_coupling(), empty_pipe(), empty_recharge(), forcing, host=host,
                     report={"workload": "plain_small", "seed": seed, "triangles": C})


def make_catchment(n_triangles: int, seed: int, spacing: float = 25.0, rain_total_mm: float = 160.0,
                   rain_hours: int = 24, interval_s: float = 3600.0, pipe_fraction: float = 0.12,
                   with_pipes: bool = True, extra_points: int = 0, relief_m: float = 120.0,
                   grid: tuple | None = None, rain_mm_h=None, horton_scale=None):
    """Seeded Delaunay catchment with fractal bed, land cover classes, Horton and pipes."""
    rng = np.random.default_rng(seed)
    if grid is None:
        side = max(int(round(np.sqrt(n_triangles / 2.0))), 2)
        nx = ny = side
    else:
        nx, ny = grid
    extent = nx * spacing
    nodes, tri = jittered_delaunay(nx, ny, extent, ny * spacing, rng, extra_points=extra_points)
    cen = cell_centroids(nodes, tri)
    bed = (relief_m * value_noise(cen[:, 0], cen[:, 1], extent, rng)).astype(np.float32)
    cover = value_noise(cen[:, 0], cen[:, 1], extent, rng, octaves=3, base_cells=8)
    cls = np.searchsorted(np.quantile(cover, [0.2, 0.4, 0.6, 0.8]), cover).astype(np.int64)
    n_by_class = np.array([0.015, 0.03, 0.05, 0.07, 0.10], np.float32)
    f0_mm = rng.uniform(0.013, 7.2, 5)
    fc_mm = np.minimum(rng.uniform(0.004, 2.25, 5), f0_mm)
    surface, tri = build_surface_topology(nodes, tri, bed, n_by_class[cls], boundary_mode="outflow")
    C = tri.shape[0]
    f0 = (f0_mm[cls] * MM_HR_TO_M_S).astype(np.float32)
    fc = (fc_mm[cls] * MM_HR_TO_M_S).astype(np.float32)
    k = np.full(C, 4.0 / 3600.0, np.float32)
    if horton_scale is not None:
        f0 = (f0 * np.float32(horton_scale[0])).astype(np.float32)
        fc = (fc * np.float32(horton_scale[1])).astype(np.float32)
        k = (k * np.float32(horton_scale[2])).astype(np.float32)
    if rain_mm_h is None:
        rain_mm_h = hourly_hyetograph(rain_total_mm, rain_hours, rng)
    forcing = rain_forcing(rain_mm_h, interval_s, C, f0=f0, fc=fc, k=k)
    if with_pipes:
        nn = max(int(round(pipe_fraction * C)), 4)
        coupling, pipe, v0 = pipe_network(rng, cen, bed, nn, int(round(1.177 * nn)),
                                          int(round(0.516 * nn)), max(1, nn // 2600), extent)
    else:
        coupling, pipe, v0 = empty_coupling(), empty_pipe(), None
    host = {"nodes": nodes, "triangles": tri, "centroids": cen}
    return _assemble(surface, coupling, pipe, empty_recharge(), forcing, node_volume=v0, host=host,
                     report={"workload": "catchment", "seed": seed, "triangles": C})


def make_production_scale(seed: int = BASE_SEED + 2, rain_total_mm: float = 160.0,
                          horton_scale=None, rain_mm_h=None):
    """Plain method, full production scale: 981,880 triangles and the full pipe network."""
    rng = np.random.default_rng(seed)
    nx = ny = 700
    spacing = 17300.0 / nx
    extent = 17300.0
    nodes, tri = jittered_delaunay(nx, ny, extent, extent, rng, extra_points=940)
    cen = cell_centroids(nodes, tri)
    bed = (120.0 * value_noise(cen[:, 0], cen[:, 1], extent, rng)).astype(np.float32)
    cover = value_noise(cen[:, 0], cen[:, 1], extent, rng, octaves=3, base_cells=8)
    cls = np.searchsorted(np.quantile(cover, [0.2, 0.4, 0.6, 0.8]), cover).astype(np.int64)
    n_by_class = np.array([0.015, 0.03, 0.05, 0.07, 0.10], np.float32)
    f0_mm = rng.uniform(0.013, 7.2, 5)
    fc_mm = np.minimum(rng.uniform(0.004, 2.25, 5), f0_mm)
    surface, tri = build_surface_topology(nodes, tri, bed, n_by_class[cls], boundary_mode="outflow")
    C = tri.shape[0]
    f0 = (f0_mm[cls] * MM_HR_TO_M_S).astype(np.float32)
    fc = (fc_mm[cls] * MM_HR_TO_M_S).astype(np.float32)
    k = np.full(C, 4.0 / 3600.0, np.float32)
    if horton_scale is not None:
        f0 = (f0 * np.float32(horton_scale[0])).astype(np.float32)
        fc = (fc * np.float32(horton_scale[1])).astype(np.float32)
        k = (k * np.float32(horton_scale[2])).astype(np.float32)
    if rain_mm_h is None:
        rain_mm_h = hourly_hyetograph(rain_total_mm, 24, rng)
    forcing = rain_forcing(rain_mm_h, 3600.0, C, f0=f0, fc=fc, k=k)
    coupling, pipe, v0 = pipe_network(rng, cen, bed, 118_768, 139_798, 61_317, 46, extent)
    host = {"nodes": nodes, "triangles": tri, "centroids": cen, "spacing_m": spacing}
    return _assemble(surface, coupling, pipe, empty_recharge(), forcing, node_volume=v0, host=host,
                     report={"workload": "production_scale", "seed": seed, "triangles": C})


def make_dam_break(nx: int = 2500, ny: int = 40, length: float = 10000.0, width: float = 200.0,
                   seed: int = BASE_SEED + 3, h_up: float = 10.0, h_down: float = 0.001):
    """Precision study: dam break in a frictionless flat channel with reflective walls."""
    rng = np.random.default_rng(seed)
    nodes, tri = jittered_delaunay(nx, ny, length, width, rng, jitter=0.25)
    cen = cell_centroids(nodes, tri)
    surface, tri = build_surface_topology(nodes, tri, np.zeros(tri.shape[0], np.float32), 0.0,
                                          boundary_mode="closed")
    C = tri.shape[0]
    h0 = np.where(cen[:, 0] < 0.5 * length, h_up, h_down).astype(np.float32)
    forcing = rain_forcing([0.0], 1.0, C)
    host = {"nodes": nodes, "triangles": tri, "centroids": cen}
    return _assemble(surface, empty_coupling(), empty_pipe(), empty_recharge(), forcing, h0=h0,
                     host=host, report={"workload": "dam_break", "seed": seed, "triangles": C})


def make_shared_mesh_ensemble(base, n_members: int = 16, seed: int = BASE_SEED + 5,
                              rain_range_mm=(40.0, 200.0)):
    """Members sharing ``base``'s mesh and network (arrays are shared, not copied)."""
    rng = np.random.default_rng(seed)
    totals = np.linspace(rain_range_mm[0], rain_range_mm[1], n_members)
    out = []
    f = base.forcing
    for m in range(n_members):
        fac = 1.0 + rng.uniform(-0.2, 0.2, 3)
        forcing = f._replace(
            rainfall_start_s=np.zeros(1, np.float32), rainfall_end_s=np.full(1, 3600.0, np.float32),
            rainfall_rate_mps=np.array([totals[m] * MM_HR_TO_M_S], np.float32),
            horton_f0_mps=(f.horton_f0_mps * np.float32(fac[0])).astype(np.float32),
            horton_fc_mps=(f.horton_fc_mps * np.float32(fac[1])).astype(np.float32),
            horton_k_per_s=(f.horton_k_per_s * np.float32(fac[2])).astype(np.float32))
        out.append(base._replace(forcing=forcing,
                                 report={**base.report, "member": m, "rain_total_mm": float(totals[m]),
                                         "horton_factors": fac.tolist()}))
    return out


def make_uneven_ensemble(n_members: int = 48, seed: int = BASE_SEED + 4, lo: float = 2.0e4, hi: float = 3.0e5,
                         spacing: float = 25.0):
    """Catchments with log-uniform triangle counts, each with its own storm and network."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n_members))).astype(int)
    totals = rng.uniform(20.0, 200.0, n_members)
    return [make_catchment(int(n), seed=seed * 1000 + i, spacing=spacing, rain_mm_h=[float(totals[i])],
                           rain_hours=1, interval_s=3600.0)
            for i, n in enumerate(sizes)]
