"""Real-asset input adapter: the reference's production asset contract and loader,
restated without JAX (NumPy arrays in the product's NamedTuples).

* ``validate_v1_parity_assets`` -- ``hybrid_flood/operational/parity_assets.py:47-138``:
  the four asset files, the topology arrays and their production shapes, the
  exterior boundary contract (34,390 code-4 edges), positive finite areas,
  non-negative capacities, the Horton/BCR arrays, and a sha256 per file.
* ``load_rainfall_csv`` -- ``hybrid_flood/operational/rainfall.py``: interval
  rows (``start_min``/``end_min`` with ``intensity_mmhr`` or ``depth_mm``) or
  timestamp rows (``timestamp``/``rainfall_mm``), with the same checks.
* ``load_parity_inputs`` -- ``hybrid_flood/operational/parity_inputs.py:130-220``:
  topology NPZ -> SurfaceTopology / CouplingTopology / PipeTopology; the recharge
  CSV (enabled rows, sources off the active mesh snapped to the nearest active
  centroid, :62-90); the external-inflow CSV (receivers snapped to active
  centroids, one column per (inflow id, cell), flows summed per time, :93-127);
  Horton fields in mm/h -> m/s; the initial node volume from the fill depth.

Paths relative to ``HF_CUPY_JAX_COMPLETED_PROJECT/reproduction/source/src/``.  The
file names, array names and CSV columns are the reference's asset format and
are kept as they are.  ``strict=False`` skips the production-shape and boundary
contract (for other meshes written in the same schema); everything else is
checked the same way.
"""

from __future__ import annotations

import csv
import hashlib
from pathlib import Path

import numpy as np

from .errors import ConfigurationError, DataFileMissingError, RainfallInputError
from .types import (CouplingTopology, LoadedParityInputs, ParityForcing, ParityState, PipeTopology,
                    RechargeTopology, SurfaceTopology)

MM_HR_TO_M_S = 1.0e-3 / 3600.0

TOPOLOGY_RELATIVE = Path("data/production_gurugram/production_topology/"
                         "d_june5_exact_production_topology_and_mesh.npz")
HORTON_RELATIVE = Path("data/production_gurugram/production_topology/horton_bcr_fields.npz")
RECHARGE_RELATIVE = Path("data/production_gurugram/drainage_june5/recharge/surface_recharge_exchange_table.csv")
BOUNDARY_INFLOW_RELATIVE = Path("data/production_gurugram/external_boundary_inflow/"
                                "gurugram_external_boundary_inflow_july2025_133mm_12h.csv")

# parity_assets.py:28-44 -- the production mesh and network
EXPECTED_SHAPES = {
    "nodes_world": (499_901, 2), "triangles": (981_880, 3), "centroids_world": (981_880, 2),
    "bed_elevation_m": (981_880,), "active_surface_mask": (981_880,), "surface_cell_area_m2": (981_880,),
    "surface_edge_left_cell": (1_490_015,), "surface_edge_right_cell": (1_490_015,),
    "surface_boundary_code": (1_490_015,), "inlet_cell_index": (61_317,), "inlet_node_index": (61_317,),
    "node_storage_area_m2": (118_768,), "node_initial_fill_depth_m": (118_768,),
    "link_from_node_index": (139_798,), "link_to_node_index": (139_798,),
    "link_flow_capacity_m3ps": (139_798,), "outfall_flow_capacity_m3ps": (118_768,),
}
PRODUCTION_BOUNDARY = {"4": 34_390}
HORTON_ARRAYS = ("horton_f0_mmhr", "horton_fc_mmhr", "horton_decay_per_hr", "building_fraction",
                 "impervious_fraction", "bcr_open_fraction")


def _sha256(path: Path) -> str:
    digest = hashlib.sha256()
    with path.open("rb") as handle:
        for block in iter(lambda: handle.read(1024 * 1024), b""):
            digest.update(block)
    return digest.hexdigest()


def _enabled(value) -> bool:
    return str(value).strip().lower() in {"1", "true", "yes", "y"}


def asset_paths(v1_root) -> dict[str, Path]:
    root = Path(v1_root).resolve()
    return {"topology": root / TOPOLOGY_RELATIVE, "horton": root / HORTON_RELATIVE,
            "recharge": root / RECHARGE_RELATIVE, "historical_boundary_inflow": root / BOUNDARY_INFLOW_RELATIVE}


def validate_v1_parity_assets(v1_root, *, strict: bool = True) -> dict:
    """parity_assets.py:47-138 -- fail fast unless the parity assets are present and consistent."""
    root = Path(v1_root).resolve()
    paths = asset_paths(root)
    missing = [str(p) for p in paths.values() if not p.is_file()]
    if missing:
        raise DataFileMissingError("Missing V1 parity assets: " + ", ".join(missing))
    with np.load(paths["topology"], allow_pickle=False) as topology:
        absent = sorted(set(EXPECTED_SHAPES).difference(topology.files))
        if absent:
            raise ConfigurationError(f"V1 topology is missing parity arrays: {absent}")
        if strict:
            wrong = {n: {"expected": s, "actual": topology[n].shape} for n, s in EXPECTED_SHAPES.items()
                     if topology[n].shape != s}
            if wrong:
                raise ConfigurationError(f"V1 topology shape mismatch: {wrong}")
        right = topology["surface_edge_right_cell"]
        codes, counts = np.unique(topology["surface_boundary_code"][right < 0], return_counts=True)
        boundary = {str(int(c)): int(n) for c, n in zip(codes, counts)}
        if strict and boundary != PRODUCTION_BOUNDARY:
            raise ConfigurationError("V1 exterior boundary contract changed; expected 34,390 outflow-only edges "
                                     f"(code 4), found {boundary}.")
        active = topology["active_surface_mask"].astype(bool)
        area = topology["surface_cell_area_m2"].astype(np.float64)
        if np.any(area <= 0) or not np.isfinite(area).all():
            raise ConfigurationError("V1 surface cell areas must be positive and finite.")
        for name in ("inlet_capture_capacity_m3ps", "link_flow_capacity_m3ps", "outfall_flow_capacity_m3ps"):
            values = topology[name]
            if np.any(~np.isfinite(values)) or np.any(values < 0):
                raise ConfigurationError(f"V1 {name} contains negative or non-finite values.")
        summary = {"triangle_count": int(topology["triangles"].shape[0]), "active_triangle_count": int(active.sum()),
                   "active_area_km2": float(area[active].sum() / 1e6), "edge_count": int(right.size),
                   "exterior_boundary_distribution": boundary,
                   "inlet_count": int(topology["inlet_cell_index"].size),
                   "node_count": int(topology["node_storage_area_m2"].size),
                   "link_count": int(topology["link_from_node_index"].size)}
        n_cells = int(topology["triangles"].shape[0])
    with np.load(paths["horton"], allow_pickle=False) as horton:
        absent = sorted(set(HORTON_ARRAYS).difference(horton.files))
        if absent:
            raise ConfigurationError(f"V1 Horton/BCR file is missing arrays: {absent}")
        if any(horton[n].shape != (n_cells,) for n in HORTON_ARRAYS):
            raise ConfigurationError("Every V1 Horton/BCR array must have one value per triangle.")
    with paths["recharge"].open(newline="", encoding="utf-8", errors="replace") as handle:
        rows = list(csv.DictReader(handle))
    return {"schema_version": 1, "ready": True, "source_root": str(root), "crs": "EPSG:32643",
            "reference_profile": "v1_exact_production_full_swe", "topology": summary,
            "recharge": {"input_rows": len(rows), "enabled_rows": sum(_enabled(r.get("hydraulic_enabled", ""))
                                                                     for r in rows)},
            "assets": {name: {"path": str(p), "sha256": _sha256(p), "size_bytes": p.stat().st_size}
                       for name, p in paths.items()}}


def load_rainfall_csv(path) -> dict:
    """rainfall.py -- the event as interval arrays: start_min, end_min, duration_min,
    intensity_mmhr, depth_mm (sorted by start), plus duration_s, total and peak."""
    import pandas as pd
    try:
        source = pd.read_csv(Path(path))
    except Exception as exc:      # noqa: BLE001 -- any reader failure is a rainfall input error
        raise RainfallInputError(f"Could not read rainfall CSV: {exc}") from exc
    if source.empty:
        raise RainfallInputError("Rainfall CSV has no data rows.")
    if {"start_min", "end_min"}.issubset(source.columns):
        frame = source.copy()
        for name in ("start_min", "end_min"):
            frame[name] = pd.to_numeric(frame[name], errors="coerce")
        frame["duration_min"] = frame["end_min"] - frame["start_min"]
        if "intensity_mmhr" in frame:
            frame["intensity_mmhr"] = pd.to_numeric(frame["intensity_mmhr"], errors="coerce")
            frame["depth_mm"] = frame["intensity_mmhr"] * frame["duration_min"] / 60.0
        elif "depth_mm" in frame:
            frame["depth_mm"] = pd.to_numeric(frame["depth_mm"], errors="coerce")
            frame["intensity_mmhr"] = frame["depth_mm"] * 60.0 / frame["duration_min"]
        else:
            raise RainfallInputError("Rainfall CSV requires intensity_mmhr or depth_mm.")
    else:
        if not {"timestamp", "rainfall_mm"}.issubset(source.columns):
            raise RainfallInputError("Rainfall CSV requires start_min/end_min with intensity_mmhr or depth_mm, "
                                     "or timestamp/rainfall_mm.")
        stamps = pd.to_datetime(source["timestamp"], errors="coerce", utc=True)
        if stamps.isna().any() or stamps.duplicated().any():
            raise RainfallInputError("Rainfall timestamps must be valid and unique.")
        order = np.argsort(stamps.to_numpy())
        stamps = stamps.iloc[order].reset_index(drop=True)
        rain = pd.to_numeric(source.iloc[order]["rainfall_mm"], errors="coerce").reset_index(drop=True)
        if len(stamps) < 2:
            raise RainfallInputError("Timestamp rainfall requires at least two rows.")
        durations = stamps.diff().dt.total_seconds().shift(-1)
        durations.iloc[-1] = durations.iloc[-2]
        start = (stamps - stamps.iloc[0]).dt.total_seconds() / 60.0
        frame = pd.DataFrame({"start_min": start, "end_min": start + durations / 60.0,
                              "duration_min": durations / 60.0, "intensity_mmhr": rain * 3600.0 / durations,
                              "depth_mm": rain})
    frame = frame.sort_values("start_min").reset_index(drop=True)
    numeric = frame[["start_min", "end_min", "duration_min", "intensity_mmhr", "depth_mm"]]
    if not np.isfinite(numeric.to_numpy(dtype=float)).all():
        raise RainfallInputError("Rainfall contains missing or non-finite values.")
    if (numeric < 0).any().any() or (frame["duration_min"] <= 0).any():
        raise RainfallInputError("Rainfall values and times must be non-negative with positive durations.")
    if (frame["start_min"].iloc[1:].to_numpy() < frame["end_min"].iloc[:-1].to_numpy()).any():
        raise RainfallInputError("Rainfall intervals overlap.")
    if frame["intensity_mmhr"].max() > 500 or frame["depth_mm"].sum() > 2_000:
        raise RainfallInputError("Rainfall magnitude is implausible; check units.")
    event = {k: frame[k].to_numpy(np.float64) for k in numeric.columns}
    event.update(duration_s=float(frame["end_min"].max() * 60.0), total_depth_mm=float(frame["depth_mm"].sum()),
                 peak_intensity_mm_hr=float(frame["intensity_mmhr"].max()))
    return event


def _snap(active, centroids):
    from scipy.spatial import cKDTree
    index = np.flatnonzero(active)
    return index, cKDTree(centroids[index])


def _load_recharge(path: Path, active, centroids):
    """parity_inputs.py:62-90 -- enabled rows; a source off the active mesh snaps to the nearest active centroid."""
    with path.open(newline="", encoding="utf-8", errors="replace") as handle:
        rows = [r for r in csv.DictReader(handle) if _enabled(r.get("hydraulic_enabled"))]
    index, tree = _snap(active, centroids)
    source, storage, capture, loss, stop = [], [], [], [], []
    remapped = 0
    for row in rows:
        cell = int(float(row.get("source_cell_index") or -1))
        if cell < 0 or cell >= len(active) or not active[cell]:
            _, nearest = tree.query([float(row["x_m"]), float(row["y_m"])])
            cell = int(index[nearest])
            remapped += 1
        source.append(cell)
        storage.append(float(row.get("storage_capacity_m3") or 0.0))
        capture.append(float(row.get("capture_capacity_m3ps") or 0.0))
        loss.append(float(row.get("recharge_capacity_m3ps") or 0.0))
        stop.append(float(row.get("stop_depth_m") or 0.02))
    recharge = RechargeTopology(np.asarray(source, np.int32), np.asarray(storage, np.float32),
                                np.asarray(capture, np.float32), np.asarray(loss, np.float32),
                                np.asarray(stop, np.float32))
    return recharge, {"row_count": len(rows), "remapped_source_count": remapped}


def _external_forcing(path: Path, active, centroids):
    """parity_inputs.py:93-127 -- receivers snapped to active centroids; one column per
    (inflow id, receiver cell); rows at the same time summed."""
    import pandas as pd
    frame = pd.read_csv(path)
    index, tree = _snap(active, centroids)
    ids = frame["inflow_id"].astype(str).to_numpy()
    x = pd.to_numeric(frame["x_utm43n_m"], errors="raise").to_numpy(float)
    y = pd.to_numeric(frame["y_utm43n_m"], errors="raise").to_numpy(float)
    distance, nearest = tree.query(np.column_stack((x, y)))
    receiver_by_row = index[nearest]
    time = pd.to_numeric(frame["time_s"], errors="raise").to_numpy(float)
    q = pd.to_numeric(frame["q_m3ps"], errors="raise").to_numpy(float)
    times = np.asarray(sorted(set(time)), dtype=np.float64)
    at = {float(v): i for i, v in enumerate(times)}
    key_of: dict[tuple[str, int], int] = {}
    receivers: list[int] = []
    columns: list[np.ndarray] = []
    for inflow_id, cell, t, flow in zip(ids, receiver_by_row, time, q):
        key = (str(inflow_id), int(cell))
        if key not in key_of:
            key_of[key] = len(receivers)
            receivers.append(int(cell))
            columns.append(np.zeros(len(times), np.float64))
        columns[key_of[key]][at[float(t)]] += float(flow)
    q_matrix = np.stack(columns, axis=1).astype(np.float32)
    report = {"receiver_count": len(receivers), "maximum_snap_m": float(np.max(distance)),
              "snap_warning_count": int(np.sum(distance > 500.0)),
              "peak_total_m3ps": float(np.max(q_matrix.sum(axis=1))),
              "hydrograph_volume_m3": float(np.trapezoid(q_matrix.sum(axis=1), times))}
    return times.astype(np.float32), np.asarray(receivers, np.int32), q_matrix, report


def load_parity_inputs(*, rainfall_csv, v1_root, strict: bool = True) -> LoadedParityInputs:
    """parity_inputs.py:130-220 -- the parity run's inputs from the V1 asset tree."""
    root = Path(v1_root).resolve()
    report = validate_v1_parity_assets(root, strict=strict)
    paths = asset_paths(root)
    with np.load(paths["topology"], allow_pickle=False) as d:
        host = {name: d[name] for name in d.files}
    active = host["active_surface_mask"].astype(bool)
    f32 = lambda k: np.asarray(host[k], np.float32)   # noqa: E731
    i32 = lambda k: np.asarray(host[k], np.int32)     # noqa: E731
    surface = SurfaceTopology(f32("bed_elevation_m"), f32("manning_n"), f32("surface_cell_area_m2"),
                              f32("surface_characteristic_length_m"), i32("surface_edge_left_cell"),
                              i32("surface_edge_right_cell"), f32("surface_edge_normal_x"),
                              f32("surface_edge_normal_y"), f32("surface_edge_length_m"),
                              i32("surface_boundary_code"), active)
    coupling = CouplingTopology(i32("inlet_cell_index"), i32("inlet_node_index"), i32("node_surface_cell_index"),
                                f32("node_invert_elevation_m"), f32("node_storage_area_m2"), f32("node_max_depth_m"),
                                f32("inlet_capture_capacity_m3ps"), f32("inlet_surcharge_capacity_m3ps"),
                                i32("surcharge_cell_index"), i32("surcharge_node_index"),
                                f32("surcharge_return_weight"),
                                bool(host["surcharge_return_weight_is_normalized"][0]))
    pipe = PipeTopology(i32("link_from_node_index"), i32("link_to_node_index"), f32("link_length_m"),
                        f32("link_area_m2"), f32("link_hydraulic_radius_m"), f32("link_manning_n"),
                        f32("link_flow_capacity_m3ps"), f32("link_invert_elevation_m"), f32("link_diameter_m"),
                        f32("link_minor_loss_coefficient"), i32("link_geometry_code"), f32("link_rect_width_m"),
                        f32("link_rect_height_m"), f32("outfall_area_m2"), f32("outfall_coefficient"),
                        f32("outfall_tailwater_head_m"), f32("outfall_flow_capacity_m3ps"))
    recharge, recharge_report = _load_recharge(paths["recharge"], active, host["centroids_world"])
    ext_t, receivers, ext_q, ext_report = _external_forcing(paths["historical_boundary_inflow"], active,
                                                            host["centroids_world"])
    event = load_rainfall_csv(rainfall_csv)
    with np.load(paths["horton"], allow_pickle=False) as horton:
        f0 = horton["horton_f0_mmhr"].astype(np.float32) * np.float32(MM_HR_TO_M_S)
        fc = horton["horton_fc_mmhr"].astype(np.float32) * np.float32(MM_HR_TO_M_S)
        decay = horton["horton_decay_per_hr"].astype(np.float32) / np.float32(3600.0)
    multiplier = np.asarray(host.get("rainfall_multiplier", np.ones(len(active), np.float32)), np.float32)
    forcing = ParityForcing(event["start_min"].astype(np.float32) * np.float32(60.0),
                            event["end_min"].astype(np.float32) * np.float32(60.0),
                            event["intensity_mmhr"].astype(np.float32) * np.float32(MM_HR_TO_M_S),
                            multiplier, f0, fc, decay, ext_t, receivers, ext_q)
    node_volume = f32("node_storage_area_m2") * f32("node_initial_fill_depth_m")
    n_cells = len(active)
    state = ParityState(np.zeros(n_cells, np.float32), np.zeros(n_cells, np.float32), np.zeros(n_cells, np.float32),
                        np.zeros(n_cells, np.float32), node_volume,
                        np.zeros(len(host["link_from_node_index"]), np.float32),
                        np.zeros(len(recharge.source_cell), np.float32))
    return LoadedParityInputs(surface, coupling, pipe, recharge, state, forcing, host,
                              {"assets": report, "recharge": recharge_report, "external_inflow": ext_report})
