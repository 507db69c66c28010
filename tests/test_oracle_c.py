"""The C restatement of the oracle (oracle/v1_parity.c, all host cores) against
the NumPy oracle it restates (the full step and the local-inertial mode) and
against the golden fixtures.

Both round every float32 operation in the same order and sum every scatter in
index order, so they differ only where libm's exp/pow/acos/sin and NumPy's
float32 routines round differently: float32 per step within 2e-6 m in depth and
1e-5 of max(V, 1 m^3) in node volume, float64 within 1e-12.  Results do not
depend on the thread count (bit-identical with 1 and 4 threads)."""

import numpy as np
import pytest

from helpers import load_golden, oracle_step, small_network_scenario, to_oracle
from oracle import v1_parity_c as oc
from oracle import v1_parity_np as orc


def _scenario(stage=False):
    inp = small_network_scenario(seed=5)
    stage_m = None
    if stage:
        s = inp.surface
        code = s.boundary_code.copy()
        b = np.flatnonzero(s.edge_right < 0)[::4]
        code[b] = 3
        stage_m = np.zeros(code.shape[0], np.float32)
        stage_m[b] = s.bed[s.edge_left[b]] + np.float32(0.2)
        inp = inp._replace(surface=s._replace(boundary_code=code))
    return inp, stage_m


@pytest.mark.parametrize("dtype,stage", [(np.float32, False), (np.float64, False), (np.float32, True),
                                         (np.float64, True)])
def test_steps_match_numpy_oracle(dtype, stage):
    inp, stage_m = _scenario(stage)
    io = to_oracle(inp, dtype)
    co = oc.COracle(*io[:4], io[5], dtype=dtype, boundary_stage_m=stage_m)
    a = b = io[4]
    t = 0.0
    tol_h = 2e-6 if dtype == np.float32 else 1e-12
    tol_v = 1e-5 if dtype == np.float32 else 1e-12
    for i in range(30):
        dt = 0.3 if i % 2 else 0.45
        rate = orc.rain_rate(io[5], t)
        q = orc.external_q(io[5], t + 0.5 * dt)
        a, da = oracle_step(io, a, dt, rain=rate, q=q, dtype=dtype, boundary_stage_m=stage_m)
        b, db = co.step(b, dt, rate, q)
        t += dt
        for f in ("h", "hu", "hv", "wet_time_s"):
            assert np.max(np.abs(getattr(a, f) - getattr(b, f))) <= tol_h, (i, f)
        for f in ("node_volume", "link_flow", "recharge_storage"):
            ra, rb = np.asarray(getattr(a, f), np.float64), np.asarray(getattr(b, f), np.float64)
            assert np.max(np.abs(ra - rb) / np.maximum(np.abs(ra), 1.0), initial=0.0) <= tol_v, (i, f)
        for k in orc.LEDGER_FIELDS:
            assert abs(da[k] - db[k]) <= (1e-5 if dtype == np.float32 else 1e-11) * max(abs(da[k]), 1e-3), (i, k)
    assert da["surface_capture"] > 0 or da["surface_surcharge"] > 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_local_inertial_steps_match_numpy_oracle(dtype):
    """The local-inertial surface mode (oracle/local_inertial_np.py) in C, coupled,
    with inactive cells and every boundary kind of the scenario; 30 steps."""
    inp = small_network_scenario(seed=13, h_init=0.2)
    io = to_oracle(inp, dtype)
    co = oc.COracle(*io[:4], io[5], dtype=dtype, surface_mode="local_inertial")
    a = b = io[4]
    t = 0.0
    tol_h = 2e-6 if dtype == np.float32 else 1e-12
    tol_v = 1e-5 if dtype == np.float32 else 1e-12
    for i in range(30):
        dt = 0.3 if i % 2 else 0.45
        rate = orc.rain_rate(io[5], t)
        q = orc.external_q(io[5], t + 0.5 * dt)
        a, da = oracle_step(io, a, dt, rain=rate, q=q, dtype=dtype, surface_mode="local_inertial")
        b, db = co.step(b, dt, rate, q)
        t += dt
        for f in ("h", "hu", "hv", "wet_time_s"):
            assert np.max(np.abs(getattr(a, f) - getattr(b, f))) <= tol_h, (i, f)
        for f in ("node_volume", "link_flow", "recharge_storage"):
            ra, rb = np.asarray(getattr(a, f), np.float64), np.asarray(getattr(b, f), np.float64)
            assert np.max(np.abs(ra - rb) / np.maximum(np.abs(ra), 1.0), initial=0.0) <= tol_v, (i, f)
        for k in orc.LEDGER_FIELDS:
            assert abs(da[k] - db[k]) <= (1e-5 if dtype == np.float32 else 1e-11) * max(abs(da[k]), 1e-3), (i, k)
    assert da["boundary_outflow"] > 0
    # the gravity-wave timestep and the policy advance
    from oracle import local_inertial_np as li
    assert co.stable_timestep(b) == li.stable_timestep(b, io[0], dtype=dtype)
    # 60 policy steps: the local-inertial step amplifies a last-bit difference about
    # 2x per step once the basin drains (float64: 1e-16 at step 40, 3e-15 at step 50,
    # 1.5e-12 at step 60), so float32 (NumPy's own float32 exp and power against libm) is
    # held to the same step count and float64 to the state
    x, cx, lx = orc.advance_to_time(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 60.0, 120.0,
                                    *io[:4], io[5], max_steps=10_000, surface_mode="local_inertial", dtype=dtype)
    y, cy, ly = co.advance_to_time(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 60.0, 120.0,
                                   max_steps=10_000)
    assert cx.global_step == cy.global_step and abs(cx.time_s - cy.time_s) <= 1e-9
    if dtype == np.float64:
        assert np.max(np.abs(x.h - y.h)) <= 1e-10


@pytest.mark.parametrize("name", ["coupled_f32", "coupled_f64", "coupled_li_f32", "coupled_li_f64"])
def test_golden_fixtures(name):
    inp, sched, expected, ledger, dtype = load_golden(name)
    io = to_oracle(inp, dtype)
    co = oc.COracle(*io[:4], io[5], dtype=dtype, surface_mode=sched["surface_mode"])
    st = io[4]
    led = orc.zero_ledger()
    for dt, rain, q in zip(sched["dt"], sched["rain"], sched["q"]):
        st, d = co.step(st, float(dt), float(rain), np.asarray(q, dtype))
        for k in led:
            led[k] += d[k]
    tol = 1e-11 if dtype == np.float64 else 2e-5
    for f in ("h", "hu", "hv", "wet_time_s"):
        np.testing.assert_allclose(getattr(st, f), getattr(expected, f), rtol=0, atol=tol, err_msg=f)
    for k in orc.LEDGER_FIELDS:
        assert abs(led[k] - ledger[k]) <= (1e-10 if dtype == np.float64 else 1e-4) * max(abs(ledger[k]), 1e-3), k


def test_stable_timestep_and_policy_advance():
    inp, _ = _scenario()
    for dtype in (np.float32, np.float64):
        io = to_oracle(inp, dtype)
        co = oc.COracle(*io[:4], io[5], dtype=dtype)
        st = io[4]._replace(hu=(io[4].h * dtype(0.3)).astype(dtype))
        assert co.stable_timestep(st) == orc.stable_timestep(st, io[0], dtype=dtype)
    io = to_oracle(inp)
    co = oc.COracle(*io[:4], io[5])
    a, ca, la = orc.advance_to_time(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 60.0, 120.0,
                                    *io[:4], io[5], max_steps=10_000)
    b, cb, lb = co.advance_to_time(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 60.0, 120.0,
                                   max_steps=10_000)
    assert ca.global_step == cb.global_step and abs(ca.time_s - cb.time_s) <= 1e-9
    assert np.max(np.abs(a.h - b.h)) <= 1e-5
    for k in ("rainfall_input", "surface_capture", "outfall_discharge"):
        assert abs(la[k] - lb[k]) <= 1e-4 * max(abs(la[k]), 1e-3), k


def test_results_do_not_depend_on_the_thread_count():
    inp, stage_m = _scenario(stage=True)
    io = to_oracle(inp)
    runs = []
    default = oc.set_threads(0)
    for mode in ("full_swe", "local_inertial"):
        for threads in (1, 4):
            oc.set_threads(threads)
            co = oc.COracle(*io[:4], io[5], boundary_stage_m=stage_m, surface_mode=mode)
            st, led = io[4], None
            for _ in range(10):
                st, led = co.step(st, 0.4, 2e-5, np.full(3, 0.2, np.float32))
            runs.append((st, led))
    oc.set_threads(default)
    for i in (0, 2):
        for f in orc.ParityState._fields:
            np.testing.assert_array_equal(getattr(runs[i][0], f), getattr(runs[i + 1][0], f), err_msg=f)
        assert runs[i][1] == runs[i + 1][1]
