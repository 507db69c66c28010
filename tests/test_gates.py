"""The eight acceptance gates (parity_validation.py:179-192) on synthetic run records."""

import numpy as np

from floodhip import gates


def _run(depth, times=(1.0, 3600.0), max_depth=None, **ledger):
    rep = {"max_depth_m": float(np.max(depth)) if max_depth is None else max_depth,
           "v1_compatible_mass_residual_ratio": 0.001, "mass_residual_ratio": 0.001}
    base = {k: 100.0 for k in gates.LEDGER_TERMS}
    base.update(ledger)
    rep.update(base)
    return {"snapshot_times_s": np.asarray(times, np.float32), "depth_snapshots_m": np.asarray(depth, np.float32),
            "report": rep}


def test_identical_runs_pass_all_gates():
    rng = np.random.default_rng(0)
    d = np.abs(rng.normal(0.1, 0.1, (2, 500)))
    res = gates.validate(_run(d), _run(d), np.full(500, 10.0))
    assert res["passed"] and all(res["gates"].values())


def test_each_gate_trips():
    rng = np.random.default_rng(1)
    d = np.abs(rng.normal(0.2, 0.1, (2, 500)))
    area = np.full(500, 10.0)
    ref = _run(d)
    bad = d.copy(); bad[0, 0] = -1e-3
    assert not gates.validate(ref, _run(bad), area)["gates"]["finite_non_negative_depth"]
    assert not gates.validate(ref, _run(d, times=(1.0, 3602.0)), area)["gates"]["snapshot_alignment_le_1s"]
    assert not gates.validate(ref, _run(d + 0.06), area)["gates"]["all_hour_depth_rmse_le_0_05m"]
    sh = d.copy(); sh[:, :100] = 0.0
    assert not gates.validate(ref, _run(sh), area)["gates"]["all_hour_csi_ge_0_95"]
    assert not gates.validate(ref, _run(d * 1.03), area)["gates"]["all_hour_volume_error_le_2pct"]
    # relative errors divide by max(|reference|, 1) (parity_validation.py:33-34)
    assert gates.validate(_run(d, max_depth=0.5), _run(d, max_depth=0.53), area)["gates"][
        "overall_max_depth_error_le_5pct"]
    assert not gates.validate(_run(d, max_depth=2.0), _run(d, max_depth=2.12), area)["gates"][
        "overall_max_depth_error_le_5pct"]
    assert not gates.validate(ref, _run(d, rainfall_input_m3=103.0), area)["gates"][
        "source_sink_ledger_error_le_2pct"]
    worse = _run(d)
    worse["report"]["v1_compatible_mass_residual_ratio"] = 0.0025
    assert not gates.validate(ref, worse, area)["gates"]["mass_residual_regression_le_0_1pct"]


def test_max_abs_depth_check():
    a = np.zeros(10)
    assert gates.max_abs_depth_check(a, a + 5e-5)["passed"]
    assert not gates.max_abs_depth_check(a, a + 2e-4)["passed"]


def test_ledger_gate_reads_v1_recharge_fields():
    """A V1 run report names the recharge terms surface_recharge_* (parity_validation.py:21-22);
    the candidate writes recharge_*.  A 3 % recharge mismatch against such a report trips gate 7."""
    rng = np.random.default_rng(2)
    d = np.abs(rng.normal(0.2, 0.1, (2, 500)))
    area = np.full(500, 10.0)
    v1 = _run(d)
    rep = v1["report"]
    for term in ("recharge_capture_m3", "recharge_loss_m3"):
        rep["surface_" + term] = rep.pop(term)
    ok = gates.validate(v1, _run(d), area)
    assert ok["gates"]["source_sink_ledger_error_le_2pct"] and ok["ledger_terms_missing"] == []
    assert ok["ledger_comparison"]["recharge_loss_m3"]["reference_field"] == "surface_recharge_loss_m3"
    for term in ("recharge_capture_m3", "recharge_loss_m3"):
        res = gates.validate(v1, _run(d, **{term: 103.0}), area)
        assert not res["gates"]["source_sink_ledger_error_le_2pct"], term
        assert abs(res["ledger_comparison"][term]["relative_error_fraction"] - 0.03) < 1e-12


def test_ledger_gate_lists_missing_terms():
    rng = np.random.default_rng(3)
    d = np.abs(rng.normal(0.2, 0.1, (2, 500)))
    ref = _run(d)
    del ref["report"]["recharge_loss_m3"]
    import pytest
    with pytest.warns(UserWarning, match="6 of 7"):
        res = gates.validate(ref, _run(d), np.full(500, 10.0))
    assert res["ledger_terms_missing"] == ["recharge_loss_m3"]
