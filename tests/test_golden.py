"""The golden fixtures (tests/golden/*.npz) against the oracle.

The fixtures hold seeded coupled scenarios and the oracle's state and summed
ledger after a stored step schedule.  Re-running the oracle on the stored inputs
must reproduce them: float64 to 1e-12 m (and 1e-12 relative in the ledger),
float32 to 2e-6 m, so a change to the oracle's arithmetic shows up here.  The
stored inputs are also checked to be what the scenario generator builds, so a
silent change to the generator is caught too.  The coupled_li_* fixtures run the
local-inertial surface step."""

import numpy as np
import pytest

from helpers import load_golden, oracle_step, small_network_scenario, to_oracle
from oracle import v1_parity_np as orc

CASES = ["coupled_f32", "coupled_f64", "coupled_li_f32", "coupled_li_f64"]


def _rerun(inp, sched, dtype):
    io = to_oracle(inp, dtype)
    st = io[4]
    ledger = orc.zero_ledger()
    for dt, rain, q in zip(sched["dt"], sched["rain"], sched["q"]):
        st, diag = oracle_step(io, st, float(dt), rain=float(rain), q=np.asarray(q, dtype), dtype=dtype,
                               surface_mode=sched["surface_mode"])
        for k in orc.LEDGER_FIELDS:
            ledger[k] += float(diag[k])
    return st, ledger


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name):
    inp, sched, expected, ledger, dtype = load_golden(name)
    st, led = _rerun(inp, sched, dtype)
    tol = 1e-12 if dtype == np.float64 else 2e-6
    for f in ("h", "hu", "hv", "wet_time_s"):
        np.testing.assert_allclose(getattr(st, f), getattr(expected, f), rtol=0, atol=tol, err_msg=f)
    for f in ("node_volume", "link_flow", "recharge_storage"):
        ref = np.asarray(getattr(expected, f), np.float64)
        np.testing.assert_allclose(getattr(st, f), ref, rtol=tol * 10, atol=tol * 10, err_msg=f)
    for k in orc.LEDGER_FIELDS:
        assert abs(led[k] - ledger[k]) <= 10 * tol * max(abs(ledger[k]), 1.0), (k, led[k], ledger[k])
    # the schedule did something: water moved, drainage was busy
    assert np.max(np.abs(expected.h - inp.initial_state.h)) > 1e-3
    assert ledger["surface_capture"] > 0 and ledger["rainfall_input"] > 0


@pytest.mark.parametrize("name,make", [
    ("coupled_f32", lambda: small_network_scenario(seed=7)),
    ("coupled_f64", lambda: small_network_scenario(seed=21, h_init=0.2, with_external=False)),
    ("coupled_li_f32", lambda: small_network_scenario(seed=13, h_init=0.2)),
    ("coupled_li_f64", lambda: small_network_scenario(seed=29, h_init=0.2, with_external=False)),
])
def test_fixture_inputs_match_generator(name, make):
    inp, *_ = load_golden(name)
    gen = make()
    for g in ("surface", "coupling", "pipe", "recharge", "initial_state", "forcing"):
        for f, v in getattr(gen, g)._asdict().items():
            np.testing.assert_array_equal(np.asarray(getattr(getattr(inp, g), f)), np.asarray(v),
                                          err_msg=f"{g}.{f}")
