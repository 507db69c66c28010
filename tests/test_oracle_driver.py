"""Pins the oracle's driver: the dt policy of choose_dt / advance_to_time
(hybrid_flood/operational/parity_runner.py:118-158) and the external-inflow
clamp of apply_external_inflow (hybrid_flood/jax_solver/v1_parity.py:407-412).

The scenario is a dry, flat, closed basin without rain or infiltration: no edge
carries flux (both sides dry), so the state never changes and the CFL time of
every evaluation is the closed form cfl * Lc / sqrt(g * min_depth) of the
smallest characteristic length.  Each step's dt then follows from the policy
alone:

* steps 0, 32, 64, ... evaluate the CFL time and use it;
* the other steps use float32(0.8 * the last evaluated value), never compounded;
* dt is capped at 1 s, truncated at final - t, floored at 1e-3 s;
* the cached value is the evaluated one, before the cap;
* the loop stops at the step budget or when t reaches the target.
"""

import math

import numpy as np

from helpers import basin, to_oracle
from oracle import v1_parity_np as orc

G, EPS, CFL = 9.80665, 1.0e-5, 0.85


def _dry_basin(evaluated_s):
    """Dry closed basin whose smallest characteristic length gives the CFL time evaluated_s."""
    io = list(to_oracle(basin(n=4, extent=40.0, code="closed")))
    s = io[0]
    lc_min = evaluated_s * math.sqrt(G * EPS) / CFL
    lc = np.full(s.characteristic_length.shape, 50.0, np.float32)
    lc[5] = np.float32(lc_min)
    io[0] = s._replace(characteristic_length=lc)
    return io


def _dt_sequence(io, n_steps, final_time_s, target_time_s=None, clock=None):
    s, c, p, r, st, f = io
    clock = clock or orc.OracleClock(0.0, 1.0, 0)
    ledger = orc.zero_ledger()
    dts, evals = [], []
    for _ in range(n_steps):
        t0 = clock.time_s
        st, clock, ledger = orc.advance_to_time(st, clock, ledger, target_time_s or 1.0e9, final_time_s,
                                                s, c, p, r, f, max_steps=1)
        if clock.time_s == t0:
            break
        dts.append(clock.time_s - t0)
        evals.append(clock.cached_dt_s)
    return np.asarray(dts), np.asarray(evals), st, clock


def test_state_is_static():
    io = _dry_basin(0.4)
    _, _, st, _ = _dt_sequence(io, 40, 1.0e6)
    np.testing.assert_array_equal(st.h, io[4].h)
    assert not np.any(st.hu) and not np.any(st.hv)


def test_evaluate_every_32nd_step_reuse_without_compounding():
    E = 0.4
    io = _dry_basin(E)
    dts, evals, _, clock = _dt_sequence(io, 70, 1.0e6)
    assert dts.size == 70 and clock.global_step == 70
    ev = dts[0]
    assert abs(ev - E) <= 2e-7 * E                               # the closed form, in float32
    reuse = float(np.float32(np.float32(ev) * np.float32(0.8)))   # 0.8 x evaluated, once
    for i, dt in enumerate(dts):
        if i % 32 == 0:
            assert dt == ev, i
        else:
            assert dt == reuse, (i, dt, reuse)
    assert np.all(evals == ev)
    assert clock.time_s == sum(float(d) for d in dts)


def test_cap_applies_after_caching_the_evaluated_time():
    # evaluated 1.1 s: the evaluating steps are capped at 1 s, the reuse steps take
    # 0.8 x 1.1 = 0.88 s (a policy that cached the capped value would give 0.8 s)
    io = _dry_basin(1.1)
    dts, evals, _, _ = _dt_sequence(io, 40, 1.0e6)
    assert dts[0] == 1.0 and dts[32] == 1.0
    assert abs(evals[0] - 1.1) <= 3e-7
    assert np.all(np.abs(dts[1:32] - 0.88) <= 1e-6)
    assert dts[1] == float(np.float32(np.float32(evals[0]) * np.float32(0.8)))
    # evaluated 2 s: the reuse value 1.6 s is capped too
    dts, _, _, _ = _dt_sequence(_dry_basin(2.0), 34, 1.0e6)
    assert np.all(dts == 1.0)


def test_truncation_at_the_event_end_and_the_floor():
    io = _dry_basin(5.0)             # every step takes the 1 s cap
    dts, _, _, clock = _dt_sequence(io, 20, 10.3, target_time_s=10.3)
    assert dts.size == 11 and np.all(dts[:10] == 1.0)
    assert abs(dts[10] - 0.3) <= 1e-5 and abs(clock.time_s - 10.3) <= 1e-5
    # 0.5 ms before the end the step is floored at 1 ms (and passes the end)
    dts, _, _, clock = _dt_sequence(io, 20, 5.0005, target_time_s=5.0005)
    assert dts.size == 6 and np.all(dts[:5] == 1.0)
    assert abs(dts[5] - 1.0e-3) <= 1e-9 and clock.time_s > 5.0005
    # past the event end every step is floored: the target, not the end, stops the loop
    dts, _, _, _ = _dt_sequence(io, 8, 5.0005, target_time_s=5.004)
    assert dts.size == 8 and np.all(np.abs(dts[5:] - 1.0e-3) <= 1e-9)


def test_target_and_step_budget_stop_the_loop():
    io = _dry_basin(5.0)
    s, c, p, r, st, f = io
    _, clock, _ = orc.advance_to_time(st, orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 1.0e9, 1.0e9,
                                      s, c, p, r, f, max_steps=70)
    assert clock.global_step == 70 and clock.time_s == 70.0
    # stops at the first step whose end reaches the target (no truncation at a target)
    _, clock, _ = orc.advance_to_time(st, orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 3.5, 1.0e9,
                                      s, c, p, r, f, max_steps=70)
    assert clock.global_step == 4 and clock.time_s == 4.0
    # a second call resumes the same policy phase from the carried clock
    _, c2, _ = orc.advance_to_time(st, clock, orc.zero_ledger(), 6.0, 1.0e9, s, c, p, r, f, max_steps=70)
    assert c2.global_step == 6


def test_negative_external_inflow_adds_nothing():
    io = _dry_basin(5.0)
    s, c, p, r, st, f = io
    rcv = np.array([3, 7, 3], np.int32)
    area = s.cell_area
    dt = np.float32(0.5)
    q = np.array([-0.5, 0.25, -2.0], np.float32)
    out, diag = orc.advance_parity_step(
        st, s, c, p, r, rainfall_rate_mps=0.0, rainfall_multiplier=f.rainfall_multiplier,
        horton_f0_mps=f.horton_f0_mps, horton_fc_mps=f.horton_fc_mps, horton_k_per_s=f.horton_k_per_s,
        external_receiver_cell=rcv, external_q_m3ps=q, dt=dt)
    assert out.h[3] == 0.0                                     # both of its rows are negative
    assert out.h[7] == np.float32(np.float32(0.25) * dt / area[7])
    assert diag["external_boundary_inflow"] == 0.25 * 0.5
    others = np.ones(area.shape, bool)
    others[[3, 7]] = False
    assert not np.any(out.h[others])


def test_external_inflow_is_sampled_at_mid_step():
    """parity_runner.py:146: q is interpolated at t + dt/2 of each step."""
    io = _dry_basin(5.0)
    s, c, p, r, st, f = io
    f = f._replace(external_times_s=np.array([0.0, 10.0], np.float32),
                   external_receiver_cell=np.array([2], np.int32),
                   external_q_m3ps=np.array([[0.0], [10.0]], np.float32))   # q(t) = t
    _, _, ledger = orc.advance_to_time(st, orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 4.0, 1.0e9,
                                       s, c, p, r, f, max_steps=10)
    # steps of 1 s at t = 0..3: q = 0.5, 1.5, 2.5, 3.5 m^3/s
    assert abs(ledger["external_boundary_inflow"] - 8.0) <= 1e-6
