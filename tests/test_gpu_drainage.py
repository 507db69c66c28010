"""GPU parity of the coupled step (surface + external inflow + inlets + links +
outfalls + overflow return + recharge) against the CPU oracle.

Tolerances (float32): per step |dh| <= 1e-5 m; node volume within 1e-4 relative
(of max(V, 1 m^3)); link flow within 1e-4 relative (of max(|Q|, 1e-3 m^3/s));
ledger terms within 1e-4 relative.  The drainage sums run in the reference's
scatter order, so differences come only from device-library transcendentals.
"""

import numpy as np
import pytest

from floodhip import synth
from floodhip.solver import HipParitySolver
from helpers import oracle_step, small_network_scenario, strip_network, to_oracle
from oracle import v1_parity_np as orc

pytestmark = pytest.mark.gpu


def _solver(inp, **kw):
    return HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing,
                           inp.initial_state, **kw)


def _compare(got, ref, diag=None, rdiag=None):
    assert np.max(np.abs(got.h - ref.h)) <= 1e-5
    vs = np.maximum(np.abs(ref.node_volume), 1.0)
    assert np.max(np.abs(got.node_volume - ref.node_volume) / vs, initial=0.0) <= 1e-4
    qs = np.maximum(np.abs(ref.link_flow), 1e-3)
    assert np.max(np.abs(got.link_flow - ref.link_flow) / qs, initial=0.0) <= 1e-4
    assert np.max(np.abs(got.recharge_storage - ref.recharge_storage), initial=0.0) <= 1e-4 * max(1.0, float(
        np.max(np.abs(ref.recharge_storage), initial=0.0)))
    if diag is not None:
        for k in orc.LEDGER_FIELDS:
            r = rdiag[k]
            assert abs(getattr(diag, k) - r) <= 1e-4 * max(abs(r), 1e-3), (k, getattr(diag, k), r)


@pytest.mark.parametrize("entries,normalized", [(True, False), (True, True), (False, False)])
def test_coupled_steps_match_oracle(entries, normalized):
    inp = small_network_scenario(with_entries=entries)
    if entries and normalized:
        w = inp.coupling.surcharge_weight
        inp = inp._replace(coupling=inp.coupling._replace(surcharge_weight=(w / w.sum()).astype(np.float32),
                                                          surcharge_weight_normalized=True))
    io = to_oracle(inp)
    st = io[4]
    t = 0.0
    with _solver(inp) as s:
        for i in range(25):
            dt = 0.3 if i % 2 else 0.45
            rate = orc.rain_rate(io[5], t)
            q = orc.external_q(io[5], t + 0.5 * dt)
            st, rdiag = oracle_step(io, st, dt, rain=rate, q=q)
            diag = s.step(dt, rate, q)
            t += dt
            got = s.get_state()
            _compare(got, st, diag, rdiag)


def test_coupled_policy_advance_matches_oracle():
    inp = small_network_scenario(seed=11, h_init=0.2)
    io = to_oracle(inp)
    st, clock, led = orc.advance_to_time(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 50.0, 120.0,
                                         *io[:4], io[5], max_steps=10_000)
    with _solver(inp) as s:
        c, ledger = s.advance_to(50.0, 120.0, max_steps=10_000)
        got = s.get_state()
    assert c["global_step"] == clock.global_step
    assert np.max(np.abs(got.h - st.h)) <= 1e-4
    for k in ("rainfall_input", "surface_capture", "outfall_discharge", "external_boundary_inflow",
              "recharge_capture"):
        r = led[k]
        assert abs(getattr(ledger, k) - r) <= 1e-3 * max(abs(r), 1e-3), (k, getattr(ledger, k), r)


def test_catchment_with_network_fixed_dt_200_steps():
    inp = synth.make_catchment(20_000, seed=synth.BASE_SEED + 40, rain_total_mm=120.0, rain_hours=1,
                               interval_s=3600.0, rain_mm_h=[120.0])
    io = to_oracle(inp)
    st = io[4]
    dt = 0.5
    for i in range(200):
        st, _ = oracle_step(io, st, dt, rain=orc.rain_rate(io[5], i * dt))
    with _solver(inp, fixed_dt=dt) as s:
        s.advance_to(200 * dt, 1e9, max_steps=200)
        got = s.get_state()
    assert np.max(np.abs(got.h - st.h)) <= 1e-4
    vs = np.maximum(np.abs(st.node_volume), 1.0)
    assert np.max(np.abs(got.node_volume - st.node_volume) / vs) <= 1e-3


def test_coupled_bitwise_repeatable():
    inp = small_network_scenario(seed=3)
    outs = []
    for _ in range(2):
        with _solver(inp, steps_per_launch=8) as s:
            s.advance_to(40.0, 100.0)
            st = s.get_state()
            outs.append(b"".join(np.asarray(x).tobytes() for x in st) + np.asarray(tuple(s.ledger())).tobytes())
    assert outs[0] == outs[1]


@pytest.mark.parametrize("make", [
    lambda: small_network_scenario(seed=21, h_init=0.2, with_external=False),
    lambda: synth.make_catchment(8000, seed=synth.BASE_SEED + 42, rain_hours=1, rain_mm_h=[150.0]),
])
def test_coupled_fp64_kernels_match_fp64_oracle(make):
    """The coupled step instantiated for double against the FP64 oracle, fixed dt,
    100 steps: both round every operation in double (FMA contraction off) and sum
    in the same orders, so depths agree to 1e-10 m; node volumes and link flows
    cross the boundary as float32 and agree to float32 rounding.  (External inflow
    is left out: the oracle interpolates the inflow table in float32.)"""
    inp = make()
    io = to_oracle(inp, np.float64)
    st = io[4]
    dt, n = 0.4, 100
    for i in range(n):
        rate = orc.rain_rate(io[5], i * dt)
        q = orc.external_q(io[5], i * dt + 0.5 * dt)
        st, _ = oracle_step(io, st, dt, rain=rate, q=q, dtype=np.float64)
    with _solver(inp, precision=64, fixed_dt=dt) as s:
        # in double, 100 x 0.4 sums to just below 40: stop on the step count
        clock, _ = s.advance_to(n * dt - 1e-6, 1e9, max_steps=n)
        h, hu, hv = s.get_state_f64()
        got = s.get_state()
    assert clock["global_step"] == n
    assert np.max(np.abs(h - st.h)) <= 1e-10
    assert np.max(np.abs(hu - st.hu)) <= 1e-10
    v_ref = st.node_volume.astype(np.float32)
    assert np.max(np.abs(got.node_volume - v_ref) / np.maximum(np.abs(v_ref), 1.0)) <= 2e-7
    q_ref = st.link_flow.astype(np.float32)
    assert np.max(np.abs(got.link_flow - q_ref) / np.maximum(np.abs(q_ref), 1e-3)) <= 2e-7


@pytest.mark.parametrize("links,inlets", [(True, False), (False, True), (True, True)])
def test_degenerate_networks_match_oracle(links, inlets):
    """Networks without pipe links, without inlets, or with neither (storage nodes
    with outfalls, overflow returns and recharge only): empty link and inlet
    tables, the kernels that serve them skipped or idle, per-step parity."""
    inp = strip_network(small_network_scenario(seed=7, h_init=0.3), links, inlets)
    io = to_oracle(inp)
    st = io[4]
    t = 0.0
    with _solver(inp) as s:
        for i in range(15):
            dt = 0.4
            rate = orc.rain_rate(io[5], t)
            q = orc.external_q(io[5], t + 0.5 * dt)
            st, rdiag = oracle_step(io, st, dt, rain=rate, q=q)
            diag = s.step(dt, rate, q)
            t += dt
            _compare(s.get_state(), st, diag, rdiag)


def test_dry_weather_steps_match_oracle_and_leave_the_run_untouched():
    """step_triangular_pipe_dry_weather_network (triangular_backend.py:833-952) on the
    GPU against the oracle's restatement: node volumes and link flows 1e-4 relative,
    the five diagnostics 1e-4 relative; the surface, the clock and the run ledger are
    unchanged, and a coupled step afterwards still matches the oracle."""
    inp = small_network_scenario(seed=7, h_init=0.3)
    io = to_oracle(inp)
    s_, c_, p_, r_, st, f_ = io
    N = c_.node_storage_area.shape[0]
    rng = np.random.default_rng(5)
    inflow = rng.uniform(-0.005, 0.03, N).astype(np.float32)
    tw = (np.asarray(p_.outfall_tailwater) + 0.05).astype(np.float32)
    with _solver(inp) as s:
        before = s.get_state()
        clock0, ledger0 = s.clock(), np.asarray(tuple(s.ledger()))
        for i, (q, t, cap) in enumerate(((inflow, None, True), (inflow, tw, True), (None, None, False))):
            st, rd = orc.dry_weather_step(st, c_, p_, 20.0, q, t, cap)
            d = s.dry_weather_step(20.0, q, t, cap)
            got = s.get_state()
            vs = np.maximum(np.abs(st.node_volume), 1.0)
            assert np.max(np.abs(got.node_volume - st.node_volume) / vs) <= 1e-4, i
            qs = np.maximum(np.abs(st.link_flow), 1e-3)
            assert np.max(np.abs(got.link_flow - st.link_flow) / qs) <= 1e-4, i
            for k, v in zip(orc.DRY_WEATHER_FIELDS, d):
                assert abs(v - rd[k]) <= 1e-4 * max(abs(rd[k]), 1e-3), (i, k, v, rd[k])
        assert np.array_equal(got.h, before.h) and np.array_equal(got.hu, before.hu)
        assert s.clock() == clock0
        assert np.array_equal(np.asarray(tuple(s.ledger())), ledger0)
        rate = orc.rain_rate(f_, 0.0)
        q = orc.external_q(f_, 0.2)
        ref, rdiag = oracle_step(io, st, 0.4, rain=rate, q=q)
        diag = s.step(0.4, rate, q)
        _compare(s.get_state(), ref, diag, rdiag)
