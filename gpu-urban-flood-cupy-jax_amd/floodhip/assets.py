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


SNAP_WARNING_M = 500.0          # external inflow points snapped farther than this are counted
RECHARGE_STOP_DEPTH_M = 0.02    # stop depth of a recharge row that leaves it blank
_TRUE_WORDS = ("1", "true", "yes", "y")


def _digest(path: Path) -> dict:
    h = hashlib.sha256()
    with path.open("rb") as f:
        while chunk := f.read(1 << 20):
            h.update(chunk)
    return {"path": str(path), "sha256": h.hexdigest(), "size_bytes": path.stat().st_size}


def _truthy(values) -> np.ndarray:
    """Element-wise: the text is one of 1/true/yes/y, ignoring case and blanks."""
    words = np.char.lower(np.char.strip(np.asarray(values, dtype=str)))
    return np.isin(words, _TRUE_WORDS)


def asset_paths(v1_root) -> dict[str, Path]:
    root = Path(v1_root).resolve()
    return {"topology": root / TOPOLOGY_RELATIVE, "horton": root / HORTON_RELATIVE,
            "recharge": root / RECHARGE_RELATIVE, "historical_boundary_inflow": root / BOUNDARY_INFLOW_RELATIVE}


def _boundary_distribution(topology) -> dict[str, int]:
    codes = topology["surface_boundary_code"][topology["surface_edge_right_cell"] < 0]
    values, counts = np.unique(codes, return_counts=True)
    return {str(int(v)): int(c) for v, c in zip(values, counts)}


def _nonnegative_finite(a) -> bool:
    a = np.asarray(a)
    return bool(np.isfinite(a).all() and not (a < 0).any())


# The topology contract of a parity run (parity_assets.py:47-104), one rule per
# row: (applies in non-strict mode, problem found in the opened archive or None).
_TOPOLOGY_RULES = (
    (True, lambda t: (f"the topology archive lacks {sorted(set(EXPECTED_SHAPES) - set(t.files))}"
                      if set(EXPECTED_SHAPES) - set(t.files) else None)),
    (False, lambda t: next((f"production shape mismatch: {n} is {t[n].shape}, the production mesh has {s}"
                            for n, s in EXPECTED_SHAPES.items() if t[n].shape != s), None)),
    (False, lambda t: (None if _boundary_distribution(t) == PRODUCTION_BOUNDARY else
                       f"exterior boundary codes {_boundary_distribution(t)} differ from the production "
                       f"contract (34,390 outflow-only code-4 edges)")),
    (True, lambda t: (None if _nonnegative_finite(t["surface_cell_area_m2"])
                      and (t["surface_cell_area_m2"] > 0).all() else "cell areas must be positive and finite")),
    (True, lambda t: next((f"{n} has negative or non-finite entries" for n in
                           ("inlet_capture_capacity_m3ps", "link_flow_capacity_m3ps", "outfall_flow_capacity_m3ps")
                           if not _nonnegative_finite(t[n])), None)),
)


def validate_v1_parity_assets(v1_root, *, strict: bool = True) -> dict:
    """The asset gate of a parity run (parity_assets.py:47-138): the four asset files
    exist; the topology archive passes _TOPOLOGY_RULES (the production shapes and
    boundary contract only when ``strict``); the Horton/BCR archive has its six
    per-triangle arrays.  Returns the readiness report: topology summary, recharge
    row counts and a sha256 per asset."""
    root = Path(v1_root).resolve()
    paths = asset_paths(root)
    absent = [str(p) for p in paths.values() if not p.is_file()]
    if absent:
        raise DataFileMissingError("V1 parity assets not found: " + ", ".join(absent))
    with np.load(paths["topology"], allow_pickle=False) as topo:
        for applies_loosely, rule in _TOPOLOGY_RULES:
            problem = rule(topo) if (strict or applies_loosely) else None
            if problem:
                raise ConfigurationError(f"V1 topology rejected: {problem}.")
        active = topo["active_surface_mask"].astype(bool)
        area = topo["surface_cell_area_m2"].astype(np.float64)
        n_cells = int(topo["triangles"].shape[0])
        summary = {"triangle_count": n_cells, "active_triangle_count": int(active.sum()),
                   "active_area_km2": float(area[active].sum() / 1e6),
                   "edge_count": int(topo["surface_edge_right_cell"].size),
                   "exterior_boundary_distribution": _boundary_distribution(topo),
                   "inlet_count": int(topo["inlet_cell_index"].size),
                   "node_count": int(topo["node_storage_area_m2"].size),
                   "link_count": int(topo["link_from_node_index"].size)}
    with np.load(paths["horton"], allow_pickle=False) as horton:
        lacking = [n for n in HORTON_ARRAYS if n not in horton.files]
        if lacking:
            raise ConfigurationError(f"V1 Horton/BCR archive lacks {lacking}.")
        if {horton[n].shape for n in HORTON_ARRAYS} != {(n_cells,)}:
            raise ConfigurationError("V1 Horton/BCR arrays must hold one value per triangle.")
    table = _read_csv_columns(paths["recharge"])
    enabled = _truthy(table.get("hydraulic_enabled", np.full(_rows(table), "")))
    return {"schema_version": 1, "ready": True, "source_root": str(root), "crs": "EPSG:32643",
            "reference_profile": "v1_exact_production_full_swe", "topology": summary,
            "recharge": {"input_rows": _rows(table), "enabled_rows": int(enabled.sum())},
            "assets": {name: _digest(p) for name, p in paths.items()}}


def _read_csv_columns(path: Path) -> dict[str, np.ndarray]:
    """Every column of a CSV as text, blanks as empty strings."""
    with path.open(newline="", encoding="utf-8", errors="replace") as f:
        reader = csv.reader(f)
        header = next(reader, [])
        body = [row + [""] * (len(header) - len(row)) for row in reader]
    return {name: np.asarray([row[i] for row in body], dtype=object).astype(str) for i, name in enumerate(header)}


def _rows(table: dict) -> int:
    return len(next(iter(table.values()))) if table else 0


def _numbers(table: dict, name: str, default: float, rows) -> np.ndarray:
    """Column ``name`` as float64 over ``rows``; blank cells and a missing column give ``default``."""
    if name not in table:
        return np.full(rows.sum(), default)
    text = table[name][rows]
    out = np.full(text.shape, default)
    given = np.char.strip(text) != ""
    out[given] = text[given].astype(np.float64)
    return out


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


class _ActiveCells:
    """Nearest active cell of a point, by cell centroid."""

    def __init__(self, active, centroids):
        from scipy.spatial import cKDTree
        self.index = np.flatnonzero(active)
        self.tree = cKDTree(np.asarray(centroids)[self.index])

    def nearest(self, xy):
        distance, k = self.tree.query(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
        return self.index[k], distance


def _load_recharge(path: Path, active, centroids):
    """Recharge rows of the exchange table (parity_inputs.py:65-91): enabled rows
    only; a source cell that is blank, off the mesh or inactive is replaced by
    the active cell whose centroid is nearest to the row's (x_m, y_m); blank
    capacities are 0 and a blank stop depth is RECHARGE_STOP_DEPTH_M."""
    table = _read_csv_columns(path)
    n = _rows(table)
    keep = _truthy(table["hydraulic_enabled"]) if "hydraulic_enabled" in table else np.zeros(n, bool)
    cell = _numbers(table, "source_cell_index", -1.0, keep).astype(np.int64)
    active = np.asarray(active, bool)
    usable = (cell >= 0) & (cell < active.size)
    usable[usable] = active[cell[usable]]
    moved = ~usable
    if moved.any():
        xy = np.column_stack((_numbers(table, "x_m", np.nan, keep)[moved], _numbers(table, "y_m", np.nan, keep)[moved]))
        cell[moved], _ = _ActiveCells(active, centroids).nearest(xy)
    recharge = RechargeTopology(cell.astype(np.int32),
                                _numbers(table, "storage_capacity_m3", 0.0, keep).astype(np.float32),
                                _numbers(table, "capture_capacity_m3ps", 0.0, keep).astype(np.float32),
                                _numbers(table, "recharge_capacity_m3ps", 0.0, keep).astype(np.float32),
                                _numbers(table, "stop_depth_m", RECHARGE_STOP_DEPTH_M, keep).astype(np.float32))
    return recharge, {"row_count": int(keep.sum()), "remapped_source_count": int(moved.sum())}


def _external_forcing(path: Path, active, centroids):
    """The external boundary hydrographs (parity_inputs.py:94-127): each row's point
    goes to its nearest active cell; every distinct (inflow id, cell) pair is one
    receiver, in order of first appearance; the table's distinct times, ascending,
    are the rows of q, and rows that share a receiver and a time add up."""
    import pandas as pd
    frame = pd.read_csv(path)
    num = {c: pd.to_numeric(frame[c], errors="raise").to_numpy(np.float64)
           for c in ("x_utm43n_m", "y_utm43n_m", "time_s", "q_m3ps")}
    cell, distance = _ActiveCells(active, centroids).nearest(np.column_stack((num["x_utm43n_m"],
                                                                              num["y_utm43n_m"])))
    keys = pd.MultiIndex.from_arrays([frame["inflow_id"].astype(str).to_numpy(), cell])
    column, receivers = pd.factorize(keys)
    times, row = np.unique(num["time_s"], return_inverse=True)
    q = np.zeros((times.size, len(receivers)))
    np.add.at(q, (row, column), num["q_m3ps"])
    q = q.astype(np.float32)
    total = q.sum(axis=1)
    report = {"receiver_count": len(receivers), "maximum_snap_m": float(distance.max()),
              "snap_warning_count": int((distance > SNAP_WARNING_M).sum()),
              "peak_total_m3ps": float(total.max()), "hydrograph_volume_m3": float(np.trapezoid(total, times))}
    return (times.astype(np.float32), np.asarray(receivers.get_level_values(1), np.int32), q, report)


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
