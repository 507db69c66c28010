"""Shared test helpers: scenario builders and oracle adapters."""

from __future__ import annotations

import numpy as np

from floodhip import synth, types
from floodhip.topology import build_surface_topology, cell_centroids
from oracle import v1_parity_np as orc


def to_oracle(inputs, dtype=np.float32):
    """Convert a LoadedParityInputs (product types) to oracle NamedTuples in `dtype`."""
    s = orc.cast_tuple(orc.SurfaceTopology(*inputs.surface), dtype)
    c = orc.cast_tuple(orc.CouplingTopology(*inputs.coupling), dtype)
    p = orc.cast_tuple(orc.PipeTopology(*inputs.pipe), dtype)
    r = orc.cast_tuple(orc.RechargeTopology(*inputs.recharge), dtype)
    st = orc.cast_tuple(orc.ParityState(*inputs.initial_state), dtype)
    f = orc.cast_tuple(orc.ParityForcing(*inputs.forcing), dtype)
    return s, c, p, r, st, f


def oracle_step(inputs_o, state, dt, rain=0.0, q=None, dtype=np.float32, surface_mode="full_swe",
                boundary_stage_m=None):
    s, c, p, r, _, f = inputs_o
    if q is None:
        q = np.zeros(np.asarray(f.external_receiver_cell).shape[0], dtype)
    return orc.advance_parity_step(
        state, s, c, p, r, rainfall_rate_mps=rain, rainfall_multiplier=f.rainfall_multiplier,
        horton_f0_mps=f.horton_f0_mps, horton_fc_mps=f.horton_fc_mps, horton_k_per_s=f.horton_k_per_s,
        external_receiver_cell=f.external_receiver_cell, external_q_m3ps=q, dt=dt, dtype=dtype,
        surface_mode=surface_mode, boundary_stage_m=boundary_stage_m)


def basin(n=12, extent=120.0, code="closed", bed=None, h0=None, manning=0.0, active=None, seed=0):
    """Split-square basin with uniform boundary code; returns LoadedParityInputs."""
    nodes, tri = synth.split_square_grid(n, extent)
    cen = cell_centroids(nodes, tri)
    z = np.zeros(tri.shape[0], np.float32) if bed is None else np.asarray(bed(cen), np.float32)
    surface, tri = build_surface_topology(nodes, tri, z, manning, active_mask=active, boundary_mode=code)
    C = tri.shape[0]
    forcing = synth.rain_forcing([0.0], 1.0e9, C)
    h = np.zeros(C, np.float32) if h0 is None else np.asarray(h0(cen), np.float32)
    state = types.ParityState(h, np.zeros(C, np.float32), np.zeros(C, np.float32), np.zeros(C, np.float32),
                              np.zeros(0, np.float32), np.zeros(0, np.float32), np.zeros(0, np.float32))
    return types.LoadedParityInputs(surface, types.empty_coupling(), types.empty_pipe(),
                                    types.empty_recharge(), state, forcing,
                                    {"nodes": nodes, "triangles": tri, "centroids": cen}, {})


def small_network_scenario(seed=7, n=24, extent=600.0, n_nodes=40, n_links=52, n_inlets=30, n_outfalls=2,
                           with_entries=True, with_recharge=True, with_external=True, h_init=0.05):
    """A small fully coupled scenario exercising every drainage path."""
    rng = np.random.default_rng(seed)
    nodes, tri = synth.jittered_delaunay(n, n, extent, extent, rng)
    cen = cell_centroids(nodes, tri)
    bed = (8.0 * synth.value_noise(cen[:, 0], cen[:, 1], extent, rng, octaves=3)).astype(np.float32)
    C = tri.shape[0]
    active = rng.random(C) > 0.05
    surface, tri = build_surface_topology(nodes, tri, bed, rng.uniform(0.02, 0.06, C).astype(np.float32),
                                          active_mask=active, boundary_mode="outflow")
    # make a few boundary edges reflective and transmissive to exercise all ghost kinds
    code = surface.boundary_code.copy()
    bidx = np.flatnonzero(surface.edge_right < 0)
    code[bidx[::5]] = 0
    code[bidx[1::7]] = 1
    code[bidx[2::11]] = 2
    surface = surface._replace(boundary_code=code)
    coupling, pipe, v0 = synth.pipe_network(rng, cen, bed, n_nodes, n_links, n_inlets, n_outfalls, extent)
    # shallow pipes so surcharge, overflow and capture all happen
    coupling = coupling._replace(node_max_depth=np.minimum(coupling.node_max_depth, 1.2).astype(np.float32),
                                 node_invert=(bed[coupling.node_surface_cell] - 1.2).astype(np.float32))
    v0 = (coupling.node_storage_area * np.float32(0.9) * coupling.node_max_depth).astype(np.float32)
    if with_entries:
        S = 25
        sc = rng.integers(0, C, S).astype(np.int32)
        sn = rng.integers(0, n_nodes // 2, S).astype(np.int32)
        sn[3] = -1  # an invalid row
        coupling = coupling._replace(surcharge_cell=sc, surcharge_node=sn,
                                     surcharge_weight=rng.uniform(0.0, 2.0, S).astype(np.float32))
    nsc = coupling.node_surface_cell.copy()
    nsc[-1] = -1   # a node without a surface cell: its overflow is unreturned
    coupling = coupling._replace(node_surface_cell=nsc)
    if with_recharge:
        R = 12
        rc = rng.integers(0, C, R).astype(np.int32)
        rc[1] = rc[0]
        recharge = types.RechargeTopology(rc, rng.uniform(5, 50, R).astype(np.float32),
                                          rng.uniform(0.01, 0.1, R).astype(np.float32),
                                          rng.uniform(0.001, 0.01, R).astype(np.float32),
                                          rng.uniform(0.0, 0.03, R).astype(np.float32))
    else:
        recharge = types.empty_recharge()
    f0 = rng.uniform(1e-6, 2e-6, C).astype(np.float32)
    fc = (f0 * 0.3).astype(np.float32)
    k = np.full(C, 4.0 / 3600.0, np.float32)
    forcing = synth.rain_forcing([30.0, 80.0, 10.0], 60.0, C, multiplier=rng.uniform(0.8, 1.2, C),
                                 f0=f0, fc=fc, k=k)
    if with_external:
        rcv = rng.integers(0, C, 3).astype(np.int32)
        rcv[2] = rcv[0]
        times = np.array([0.0, 30.0, 90.0, 400.0], np.float32)
        q = rng.uniform(0.0, 0.5, (4, 3)).astype(np.float32)
        q[1, 1] = -0.2
        forcing = forcing._replace(external_times_s=times, external_receiver_cell=rcv, external_q_m3ps=q)
    h0 = np.where(surface.active, h_init * rng.random(C), 0).astype(np.float32)
    state = types.ParityState(h0, np.zeros(C, np.float32), np.zeros(C, np.float32),
                              np.zeros(C, np.float32), v0, np.zeros(pipe.link_from.shape[0], np.float32),
                              np.zeros(recharge.source_cell.shape[0], np.float32))
    return types.LoadedParityInputs(surface, coupling, pipe, recharge, state, forcing,
                                    {"nodes": nodes, "triangles": tri, "centroids": cen}, {})


def stoker_depth(x, t, h_l, h_r, x0, g=9.80665):
    """Stoker's analytic wet-bed dam-break depth (1-D), solved for the shock state."""
    from scipy.optimize import brentq
    c_l, c_r = np.sqrt(g * h_l), np.sqrt(g * h_r)

    def f(hm):  # continuity of velocity across rarefaction and shock
        um_raref = 2.0 * (c_l - np.sqrt(g * hm))
        s = np.sqrt(0.5 * g * hm * (hm + h_r) / h_r) if h_r > 0 else 0.0  # shock speed (u_r = 0)
        um_shock = s * (1.0 - h_r / hm)
        return um_raref - um_shock

    hm = brentq(f, h_r * (1 + 1e-12), h_l)
    um = 2.0 * (c_l - np.sqrt(g * hm))
    s = np.sqrt(0.5 * g * hm * (hm + h_r) / h_r)
    xi = (x - x0) / max(t, 1e-12)
    h = np.where(xi <= -c_l, h_l, 0.0)
    raref = (xi > -c_l) & (xi <= um - np.sqrt(g * hm))
    h = np.where(raref, (2 * c_l - xi) ** 2 / (9 * g), h)
    h = np.where((xi > um - np.sqrt(g * hm)) & (xi <= s), hm, h)
    h = np.where(xi > s, h_r, h)
    return h


def oracle_run(inputs, duration_s, snapshot_interval_s=3600.0, max_steps_per_hour=20_000,
               surface_mode="full_swe"):
    """The oracle's whole-event run in the layout of ``run_hip_v1_parity``: snapshots
    at the same output targets (the first at 1e-6 s), FP64 ledger, the report fields
    the gates read (parity_runner.py:264-361 driver, v1-compatible residual)."""
    from floodhip.runner import output_targets
    s, c, p, r, st, f = to_oracle(inputs)
    area = np.asarray(inputs.surface.cell_area, np.float64)
    clock = orc.OracleClock(0.0, 1.0, 0)
    ledger = orc.zero_ledger()
    depth, times, nodes, speed, steps = [], [], [], [], []
    for i, req in enumerate(output_targets(duration_s, snapshot_interval_s)):
        target = np.float32(1.0e-6) if i == 0 else req
        st, clock, ledger = orc.advance_to_time(st, clock, ledger, float(target), float(duration_s), s, c, p, r, f,
                                                max_steps=max_steps_per_hour, surface_mode=surface_mode)
        depth.append(np.asarray(st.h, np.float32))
        nodes.append(np.asarray(st.node_volume, np.float32))
        times.append(clock.time_s)
        steps.append(clock.global_step)
        wet = st.h > 1.0e-4
        sp = np.zeros(st.h.shape, np.float32)
        sp[wet] = np.sqrt(st.hu[wet] ** 2 + st.hv[wet] ** 2) / st.h[wet]
        speed.append(sp)
    depth_a = np.stack(depth)
    init_s = float(np.sum(np.asarray(inputs.initial_state.h, np.float64) * area))
    init_n = float(np.sum(np.asarray(inputs.initial_state.node_volume, np.float64)))
    fin_s = float(np.sum(depth_a[-1].astype(np.float64) * area))
    fin_n = float(np.sum(nodes[-1].astype(np.float64)))
    lhs = init_s + init_n + ledger["rainfall_input"] + ledger["external_boundary_inflow"]
    rhs = (fin_s + fin_n + ledger["infiltration_loss"] + ledger["outfall_discharge"]
           + ledger["unreturned_overflow"] + ledger["recharge_loss"] - ledger["boundary_net_flux"])
    report = {"max_depth_m": float(np.max(depth_a)), **{f"{k}_m3": v for k, v in ledger.items()},
              "mass_residual_ratio": (lhs - rhs) / max(abs(lhs), 1.0), "step_count": clock.global_step}
    return {"snapshot_times_s": np.asarray(times, np.float32), "depth_snapshots_m": depth_a, "report": report,
            "steps_at_snapshot": np.asarray(steps, np.int64), "speed_snapshots_mps": np.stack(speed),
            "drain_node_volume_snapshots_m3": np.stack(nodes)}


GOLDEN_DIR = __import__("pathlib").Path(__file__).resolve().parent / "golden"
_GROUP_TYPES = {"surface": types.SurfaceTopology, "coupling": types.CouplingTopology, "pipe": types.PipeTopology,
                "recharge": types.RechargeTopology, "initial_state": types.ParityState,
                "forcing": types.ParityForcing}


def load_golden(name):
    """A fixture written by tests/golden/make_golden.py: (LoadedParityInputs, schedule
    dict with dt/rain/q rows and the surface mode, expected ParityState, expected
    ledger dict, dtype)."""
    with np.load(GOLDEN_DIR / f"{name}.npz", allow_pickle=False) as z:
        a = {k: z[k] for k in z.files}
    groups = {}
    for g, cls in _GROUP_TYPES.items():
        groups[g] = cls(*[(a[f"{g}.{f}"].item() if a[f"{g}.{f}"].ndim == 0 else a[f"{g}.{f}"])
                          for f in cls._fields])
    inp = types.LoadedParityInputs(groups["surface"], groups["coupling"], groups["pipe"], groups["recharge"],
                                   groups["initial_state"], groups["forcing"], {}, {})
    expected = types.ParityState(*[a[f"out.{f}"] for f in types.ParityState._fields])
    ledger = {k: float(a[f"ledger.{k}"]) for k in orc.LEDGER_FIELDS}
    mode = str(a["surface_mode"]) if "surface_mode" in a else "full_swe"
    sched = {"dt": a["dt"], "rain": a["rain"], "q": a["q"], "surface_mode": mode}
    return inp, sched, expected, ledger, np.dtype(str(a["dtype"]))


def strip_network(inp, links=True, inlets=True):
    """The scenario with its pipe links and/or its inlets removed (nodes, outfalls,
    overflow returns and recharge stay): the degenerate networks of a drainage
    model that has storage nodes without pipes or pipes without inlets."""
    coupling, pipe, state = inp.coupling, inp.pipe, inp.initial_state
    if links:
        pipe = pipe._replace(**{f: np.asarray(v)[:0] for f, v in pipe._asdict().items() if f.startswith("link_")})
        state = state._replace(link_flow=np.zeros(0, np.float32))
    if inlets:
        coupling = coupling._replace(**{f: np.asarray(v)[:0] for f, v in coupling._asdict().items()
                                        if f.startswith("inlet_")})
    return inp._replace(coupling=coupling, pipe=pipe, initial_state=state)


def write_asset_tree(root, inp, with_multiplier=True):
    """Writes ``inp`` as a V1 asset tree in the reference's schema (parity_assets.py,
    parity_inputs.py) under ``root`` and returns (topology arrays, Horton arrays,
    rainfall CSV path, an inactive cell)."""
    import csv
    from floodhip import assets
    s, c, p = inp.surface, inp.coupling, inp.pipe
    cen = np.asarray(inp.host["centroids"], np.float64)
    C = s.bed.shape[0]
    N = c.node_storage_area.shape[0]
    topo = {
        "nodes_world": np.asarray(inp.host["nodes"], np.float64), "triangles": np.asarray(inp.host["triangles"]),
        "centroids_world": cen, "bed_elevation_m": s.bed, "manning_n": s.manning_n,
        "active_surface_mask": s.active.astype(np.uint8), "surface_cell_area_m2": s.cell_area,
        "surface_characteristic_length_m": s.characteristic_length, "surface_edge_left_cell": s.edge_left,
        "surface_edge_right_cell": s.edge_right, "surface_edge_normal_x": s.edge_nx,
        "surface_edge_normal_y": s.edge_ny, "surface_edge_length_m": s.edge_length,
        "surface_boundary_code": s.boundary_code, "inlet_cell_index": c.inlet_cell, "inlet_node_index": c.inlet_node,
        "node_surface_cell_index": c.node_surface_cell, "node_invert_elevation_m": c.node_invert,
        "node_storage_area_m2": c.node_storage_area, "node_max_depth_m": c.node_max_depth,
        "inlet_capture_capacity_m3ps": c.inlet_capture_capacity,
        "inlet_surcharge_capacity_m3ps": c.inlet_surcharge_capacity, "surcharge_cell_index": c.surcharge_cell,
        "surcharge_node_index": c.surcharge_node, "surcharge_return_weight": c.surcharge_weight,
        "surcharge_return_weight_is_normalized": np.array([c.surcharge_weight_normalized]),
        "link_from_node_index": p.link_from, "link_to_node_index": p.link_to, "link_length_m": p.link_length,
        "link_area_m2": p.link_area, "link_hydraulic_radius_m": p.link_radius, "link_manning_n": p.link_manning,
        "link_flow_capacity_m3ps": p.link_capacity, "link_invert_elevation_m": p.link_invert,
        "link_diameter_m": p.link_diameter, "link_minor_loss_coefficient": p.link_minor_loss,
        "link_geometry_code": p.link_geometry_code, "link_rect_width_m": p.link_rect_width,
        "link_rect_height_m": p.link_rect_height, "outfall_area_m2": p.outfall_area,
        "outfall_coefficient": p.outfall_coefficient, "outfall_tailwater_head_m": p.outfall_tailwater,
        "outfall_flow_capacity_m3ps": p.outfall_capacity,
        "node_initial_fill_depth_m": np.linspace(0.0, 0.5, N).astype(np.float32),
    }
    if with_multiplier:
        topo["rainfall_multiplier"] = np.linspace(0.8, 1.2, C).astype(np.float32)
    paths = assets.asset_paths(root)
    for path in paths.values():
        path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(paths["topology"], **topo)
    rng = np.random.default_rng(0)
    horton = {"horton_f0_mmhr": rng.uniform(5, 20, C).astype(np.float32),
              "horton_fc_mmhr": rng.uniform(1, 5, C).astype(np.float32),
              "horton_decay_per_hr": rng.uniform(1, 4, C).astype(np.float32)}
    horton.update({k: np.zeros(C, np.float32) for k in ("building_fraction", "impervious_fraction",
                                                         "bcr_open_fraction")})
    np.savez(paths["horton"], **horton)
    inactive = int(np.flatnonzero(~s.active)[0])
    with paths["recharge"].open("w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["hydraulic_enabled", "source_cell_index", "x_m", "y_m", "storage_capacity_m3",
                    "capture_capacity_m3ps", "recharge_capacity_m3ps", "stop_depth_m"])
        w.writerow(["true", 5, 0, 0, 10.0, 0.05, 0.005, 0.03])
        w.writerow(["no", 6, 0, 0, 99.0, 0.05, 0.005, 0.03])                 # disabled
        w.writerow(["1", inactive, cen[inactive, 0], cen[inactive, 1], 20.0, 0.02, 0.001, ""])   # off the mesh
    with paths["historical_boundary_inflow"].open("w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["inflow_id", "x_utm43n_m", "y_utm43n_m", "time_s", "q_m3ps"])
        for t, q in ((0, 0.1), (600, 0.5), (1200, 0.2)):
            w.writerow(["A", cen[10, 0], cen[10, 1], t, q])
            w.writerow(["B", cen[40, 0], cen[40, 1], t, 2 * q])
        w.writerow(["A", cen[10, 0], cen[10, 1], 600, 0.25])                 # same id, cell and time: summed
    rain = root / "rain.csv"
    with rain.open("w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["start_min", "end_min", "intensity_mmhr"])
        w.writerows([(0, 30, 40.0), (30, 60, 90.0), (60, 120, 10.0)])
    return topo, horton, rain, inactive
