"""The reference's eight cross-backend acceptance gates.

Restates ``hybrid_flood/operational/parity_validation.py:95-192`` over
in-memory run results (``run_hip_v1_parity`` returns the same fields the
reference writes to ``flood_simulation_outputs.npz`` and ``run_report.json``),
plus ``validate_run_dirs`` for two directories in the reference's on-disk
layout.  ``max_abs_depth_check`` is the additional fixed-dt check (max |dh|).
"""

from __future__ import annotations

import json
import warnings
from pathlib import Path

import numpy as np

from .errors import ConfigurationError

DEPTH_THRESHOLD_M = 0.05
# candidate report field -> reference report field (parity_validation.py:15-23).  The
# reference side of a gate check is a V1 run report, which names the two recharge
# terms surface_recharge_*; the candidate (this backend, like the JAX runner) writes
# recharge_*.
LEDGER_TERMS = {
    "rainfall_input_m3": "rainfall_input_m3",
    "infiltration_loss_m3": "infiltration_loss_m3",
    "boundary_net_flux_m3": "boundary_net_flux_m3",
    "external_boundary_inflow_m3": "external_boundary_inflow_m3",
    "outfall_discharge_m3": "outfall_discharge_m3",
    "recharge_capture_m3": "surface_recharge_capture_m3",
    "recharge_loss_m3": "surface_recharge_loss_m3",
}


def _reference_term(report: dict, cand_term: str) -> str | None:
    """The reference report's field for one candidate ledger term: the V1 name, or,
    for a reference written in the candidate's own layout (an oracle run or another
    run of this backend), the candidate's name.  None when the report has neither."""
    v1 = LEDGER_TERMS[cand_term]
    if v1 in report:
        return v1
    return cand_term if cand_term in report else None


def _relative(candidate: float, reference: float) -> float:
    """parity_validation.py:33-34."""
    return (candidate - reference) / max(abs(reference), 1.0)


def validate(reference: dict, candidate: dict, area: np.ndarray, valid: np.ndarray | None = None,
             threshold_m: float = DEPTH_THRESHOLD_M) -> dict:
    """Apply the eight gates.  ``reference``/``candidate`` carry ``snapshot_times_s``,
    ``depth_snapshots_m`` and a ``report`` dict with ``max_depth_m``, the ledger
    terms and ``mass_residual_ratio`` (reference) / ``v1_compatible_mass_residual_ratio``."""
    ref_times = np.asarray(reference["snapshot_times_s"], np.float64)
    cand_times = np.asarray(candidate["snapshot_times_s"], np.float64)
    ref_depth = np.asarray(reference["depth_snapshots_m"], np.float32)
    cand_depth = np.asarray(candidate["depth_snapshots_m"], np.float32)
    if ref_depth.shape != cand_depth.shape or ref_times.shape != cand_times.shape:
        raise ValueError(f"Parity shape mismatch: reference {ref_depth.shape}/{ref_times.shape}, "
                         f"candidate {cand_depth.shape}/{cand_times.shape}.")
    area = np.asarray(area, np.float64)
    valid = np.ones(area.shape, bool) if valid is None else np.asarray(valid, bool)
    finite_non_negative = bool(np.isfinite(cand_depth).all() and np.min(cand_depth) >= -1.0e-7)
    hourly = []
    for i, (rt, ct) in enumerate(zip(ref_times, cand_times)):
        rv = ref_depth[i, valid].astype(np.float64)
        cv = cand_depth[i, valid].astype(np.float64)
        diff = cv - rv
        ref_wet, cand_wet = rv >= threshold_m, cv >= threshold_m
        union, inter = ref_wet | cand_wet, ref_wet & cand_wet
        rvol = float(np.sum(ref_depth[i].astype(np.float64) * area))
        cvol = float(np.sum(cand_depth[i].astype(np.float64) * area))
        hourly.append({
            "snapshot_index": i, "reference_time_s": float(rt), "candidate_time_s": float(ct),
            "time_delta_s": float(ct - rt),
            "valid_depth_rmse_m": float(np.sqrt(np.mean(diff ** 2))),
            "depth_absolute_error_p99_m": float(np.percentile(np.abs(diff), 99)),
            "flood_extent_csi": float(inter.sum() / max(int(union.sum()), 1)),
            "reference_surface_volume_m3": rvol, "candidate_surface_volume_m3": cvol,
            "surface_volume_error_fraction": _relative(cvol, rvol),
            "reference_max_depth_m": float(np.max(rv)), "candidate_max_depth_m": float(np.max(cv)),
        })
    rrep, crep = reference["report"], candidate["report"]
    ledger, missing = {}, []
    for cand_term in LEDGER_TERMS:
        ref_term = _reference_term(rrep, cand_term)
        if ref_term is None or cand_term not in crep:
            # the reference validator skips a term absent from either report; the
            # skip is listed (and warned about) so a narrower gate 7 is visible
            missing.append(cand_term)
            continue
        ledger[cand_term] = {"reference_field": ref_term, "reference": float(rrep[ref_term]),
                             "candidate": float(crep[cand_term]),
                             "relative_error_fraction": _relative(float(crep[cand_term]), float(rrep[ref_term]))}
    if missing:
        warnings.warn(f"ledger gate compares {len(ledger)} of {len(LEDGER_TERMS)} terms; missing from a report: "
                      f"{', '.join(missing)}", stacklevel=2)
    max_dt = max(abs(r["time_delta_s"]) for r in hourly)
    max_rmse = max(r["valid_depth_rmse_m"] for r in hourly)
    min_csi = min(r["flood_extent_csi"] for r in hourly)
    max_vol = max(abs(r["surface_volume_error_fraction"]) for r in hourly)
    max_ledger = max((abs(v["relative_error_fraction"]) for v in ledger.values()), default=float("inf"))
    ref_ratio = abs(float(rrep.get("mass_residual_ratio", rrep.get("v1_compatible_mass_residual_ratio",
                                                                    float("inf")))))
    cand_ratio = abs(float(crep.get("v1_compatible_mass_residual_ratio", float("inf"))))
    gates = {
        "finite_non_negative_depth": finite_non_negative,
        "snapshot_alignment_le_1s": max_dt <= 1.0,
        "all_hour_depth_rmse_le_0_05m": max_rmse <= 0.05,
        "all_hour_csi_ge_0_95": min_csi >= 0.95,
        "all_hour_volume_error_le_2pct": max_vol <= 0.02,
        "overall_max_depth_error_le_5pct": abs(_relative(float(crep["max_depth_m"]),
                                                         float(rrep["max_depth_m"]))) <= 0.05,
        "source_sink_ledger_error_le_2pct": max_ledger <= 0.02,
        "mass_residual_regression_le_0_1pct": (cand_ratio - ref_ratio) <= 0.001,
    }
    return {
        "passed": all(gates.values()), "gates": gates,
        "summary": {"maximum_snapshot_time_delta_s": max_dt, "maximum_hourly_valid_depth_rmse_m": max_rmse,
                    "minimum_hourly_flood_extent_csi": min_csi,
                    "maximum_hourly_surface_volume_error_fraction": max_vol,
                    "maximum_source_sink_ledger_error_fraction": max_ledger,
                    "reference_mass_residual_ratio": ref_ratio,
                    "candidate_v1_compatible_mass_residual_ratio": cand_ratio},
        "hourly": hourly, "ledger_comparison": ledger, "ledger_terms_missing": missing,
    }


def max_abs_depth_check(reference_depth, candidate_depth, tolerance_m=1.0e-4) -> dict:
    err = float(np.max(np.abs(np.asarray(candidate_depth, np.float64) - np.asarray(reference_depth, np.float64))))
    return {"max_abs_depth_error_m": err, "passed": err <= tolerance_m, "tolerance_m": tolerance_m}


def _load_dir(directory: Path) -> dict:
    with np.load(directory / "flood_simulation_outputs.npz", allow_pickle=False) as d:
        out = {k: d[k] for k in d.files}
    out["report"] = json.loads((directory / "run_report.json").read_text(encoding="utf-8"))
    return out


def validate_run_dirs(reference_dir, candidate_dir, area=None, threshold_m=DEPTH_THRESHOLD_M) -> dict:
    """Gates over two run directories in the reference's layout (parity_validation.py:43-93).

    Checks what the reference's validator checks before the gates: both archives,
    snapshot times and depths of the same shapes, the candidate's active mask
    (minus its building mask) and centroids.  Without ``area`` the cell areas
    come from the topology asset named in the candidate report's input report,
    as the reference recovers them."""
    ref_dir, cand_dir = Path(reference_dir), Path(candidate_dir)
    for d in (ref_dir, cand_dir):
        if not (d / "flood_simulation_outputs.npz").is_file():
            raise ConfigurationError("Reference and candidate must both contain flood_simulation_outputs.npz.")
    ref = _load_dir(ref_dir)
    cand = _load_dir(cand_dir)
    for run in (ref, cand):
        if not {"snapshot_times_s", "depth_snapshots_m"}.issubset(run):
            raise ConfigurationError("Parity archives require snapshot_times_s and depth_snapshots_m.")
    if ref["depth_snapshots_m"].shape != cand["depth_snapshots_m"].shape or \
            ref["snapshot_times_s"].shape != cand["snapshot_times_s"].shape:
        raise ConfigurationError("Parity shape mismatch between reference and candidate.")
    if "active_surface_mask" not in cand:
        raise ConfigurationError("Candidate archive lacks active_surface_mask.")
    valid = cand["active_surface_mask"].astype(bool)
    if "building_mask" in cand:
        valid = valid & ~cand["building_mask"].astype(bool)
    if valid.shape != ref["depth_snapshots_m"].shape[1:]:
        raise ConfigurationError("Candidate domain mask does not match triangle snapshots.")
    if "centroids_world" not in cand:
        raise ConfigurationError("Candidate archive lacks exact-topology centroids_world.")
    if area is None:
        topology = Path(cand["report"]["input_report"]["assets"]["assets"]["topology"]["path"])
        with np.load(topology, allow_pickle=False) as t:
            area = t["surface_cell_area_m2"].astype(np.float64)
    if np.asarray(area).shape != valid.shape:
        raise ConfigurationError("Production triangle areas do not match parity output.")
    return validate(ref, cand, area, valid, threshold_m)
