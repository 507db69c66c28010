"""The real-asset input adapter (floodhip.assets) against a synthetic asset tree
written in the reference's schema (parity_assets.py, parity_inputs.py:62-220,
rainfall.py): every array lands in its NamedTuple field, recharge sources off the
active mesh snap to the nearest active centroid, external inflow rows merge per
(inflow id, receiver cell) and sum per time, units convert, and the contract
checks raise the reference's error types."""

import csv

import numpy as np
import pytest

from floodhip import assets
from floodhip.errors import ConfigurationError, DataFileMissingError, RainfallInputError
from helpers import small_network_scenario, write_asset_tree


def test_loader_maps_every_array_and_converts_units(tmp_path):
    inp = small_network_scenario(seed=7)
    topo, horton, rain, inactive = write_asset_tree(tmp_path, inp)
    got = assets.load_parity_inputs(rainfall_csv=rain, v1_root=tmp_path, strict=False)
    for name, field in (("surface", "bed"), ("surface", "edge_right"), ("coupling", "inlet_node"),
                        ("coupling", "surcharge_weight"), ("pipe", "link_diameter"), ("pipe", "outfall_capacity")):
        np.testing.assert_array_equal(getattr(getattr(got, name), field), getattr(getattr(inp, name), field))
    assert got.coupling.surcharge_weight_normalized == inp.coupling.surcharge_weight_normalized
    f = got.forcing
    np.testing.assert_array_equal(f.rainfall_start_s, np.array([0, 1800, 3600], np.float32))
    np.testing.assert_array_equal(f.rainfall_end_s, np.array([1800, 3600, 7200], np.float32))
    np.testing.assert_allclose(f.rainfall_rate_mps, np.array([40.0, 90.0, 10.0]) / 3.6e6, rtol=1e-6)
    np.testing.assert_allclose(f.horton_f0_mps, horton["horton_f0_mmhr"] / 3.6e6, rtol=1e-6)
    np.testing.assert_allclose(f.horton_k_per_s, horton["horton_decay_per_hr"] / 3600.0, rtol=1e-6)
    np.testing.assert_array_equal(f.rainfall_multiplier, topo["rainfall_multiplier"])
    # recharge: the disabled row is dropped, the inactive source snaps to an active cell
    r = got.recharge
    assert r.source_cell.shape == (2,) and r.source_cell[0] == 5
    assert inp.surface.active[r.source_cell[1]] and r.source_cell[1] != inactive
    np.testing.assert_allclose(r.stop_depth, [0.03, 0.02], rtol=1e-6)
    assert got.report["recharge"] == {"row_count": 2, "remapped_source_count": 1}
    # external inflow: one column per (id, cell), the repeated (A, 600 s) row summed
    np.testing.assert_array_equal(f.external_times_s, np.array([0, 600, 1200], np.float32))
    assert f.external_receiver_cell.tolist() == [10, 40]
    np.testing.assert_allclose(f.external_q_m3ps, [[0.1, 0.2], [0.75, 1.0], [0.2, 0.4]], rtol=1e-6)
    # initial node volume from the fill depth; everything else starts dry
    np.testing.assert_allclose(got.initial_state.node_volume,
                               topo["node_storage_area_m2"] * topo["node_initial_fill_depth_m"], rtol=1e-6)
    assert not np.any(got.initial_state.h)
    assert got.report["assets"]["topology"]["inlet_count"] == inp.coupling.inlet_cell.shape[0]
    assert len(got.report["assets"]["assets"]["topology"]["sha256"]) == 64


def test_strict_contract_and_missing_assets(tmp_path):
    inp = small_network_scenario(seed=7)
    _, _, rain, _ = write_asset_tree(tmp_path, inp)
    with pytest.raises(ConfigurationError, match="shape mismatch"):
        assets.load_parity_inputs(rainfall_csv=rain, v1_root=tmp_path, strict=True)
    assets.asset_paths(tmp_path)["horton"].unlink()
    with pytest.raises(DataFileMissingError):
        assets.validate_v1_parity_assets(tmp_path, strict=False)


def test_rainfall_rules(tmp_path):
    def write(name, header, rows):
        p = tmp_path / name
        with p.open("w", newline="") as f:
            w = csv.writer(f)
            w.writerow(header)
            w.writerows(rows)
        return p
    ev = assets.load_rainfall_csv(write("d.csv", ["start_min", "end_min", "depth_mm"], [(0, 30, 10), (30, 90, 6)]))
    np.testing.assert_allclose(ev["intensity_mmhr"], [20.0, 6.0])
    assert ev["duration_s"] == 5400.0 and ev["total_depth_mm"] == 16.0
    ts = assets.load_rainfall_csv(write("t.csv", ["timestamp", "rainfall_mm"],
                                        [("2025-07-01T00:10:00Z", 3), ("2025-07-01T00:00:00Z", 5)]))
    np.testing.assert_allclose(ts["start_min"], [0.0, 10.0])
    np.testing.assert_allclose(ts["intensity_mmhr"], [30.0, 18.0])     # the last row reuses the previous duration
    for rows, msg in (([(0, 30, 10), (20, 60, 5)], "overlap"), ([(0, 30, 600)], "implausible"),
                      ([(10, 10, 5)], "positive durations")):
        with pytest.raises(RainfallInputError, match=msg):
            assets.load_rainfall_csv(write("bad.csv", ["start_min", "end_min", "intensity_mmhr"], rows))
    with pytest.raises(RainfallInputError):
        assets.load_rainfall_csv(write("none.csv", ["a", "b"], [(1, 2)]))


def test_published_runs_validate_like_the_reference(tmp_path):
    """Adapter -> oracle run -> publish -> validate_run_dirs: the archive and report
    carry what the reference's validator reads (parity_validation.py:43-93), and
    without explicit areas the validator finds them through the candidate report's
    topology asset, as the reference does."""
    import json
    from floodhip import gates, runner
    from helpers import oracle_run
    inp = small_network_scenario(seed=7, h_init=0.15)
    _, _, rain, _ = write_asset_tree(tmp_path / "v1", inp)
    loaded = assets.load_parity_inputs(rainfall_csv=rain, v1_root=tmp_path / "v1", strict=False)
    # start wet: the CSI gate is 0 on a snapshot where no cell is above 5 cm
    h0 = np.where(loaded.surface.active, 0.1, 0.0).astype(np.float32)
    loaded = loaded._replace(initial_state=loaded.initial_state._replace(h=h0))
    res = oracle_run(loaded, 60.0, snapshot_interval_s=30.0)
    res["report"]["input_report"] = loaded.report
    res["report"]["v1_compatible_mass_residual_ratio"] = res["report"]["mass_residual_ratio"]
    ref_dir = runner.publish(res, tmp_path / "ref", loaded)
    cand_dir = runner.publish(res, tmp_path / "cand", loaded)
    with np.load(cand_dir / "flood_simulation_outputs.npz", allow_pickle=False) as d:
        for k in ("snapshot_times_s", "depth_snapshots_m", "active_surface_mask", "building_mask", "centroids_world",
                  "triangles", "nodes_world", "drain_link_from", "final_node_volume_m3", "max_speed_mps"):
            assert k in d.files, k
        assert d["depth_snapshots_m"].shape == (3, inp.surface.bed.shape[0])   # t = 1e-6, 30, 60 s
    rep = json.loads((cand_dir / "run_report.json").read_text())
    assert rep["input_report"]["assets"]["assets"]["topology"]["path"].endswith(".npz")
    out = gates.validate_run_dirs(ref_dir, cand_dir)
    assert out["passed"], out["gates"]
    with pytest.raises(ConfigurationError):
        runner.publish(res, tmp_path / "cand", loaded)        # never overwrites a run
    (tmp_path / "bad").mkdir()
    with pytest.raises(ConfigurationError):
        gates.validate_run_dirs(ref_dir, tmp_path / "bad")
