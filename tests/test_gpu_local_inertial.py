"""GPU parity of the local-inertial surface mode (surface_mode="local_inertial":
triangular_backend.py:1181-1291 with the gravity-wave CFL of engine.py:220-225)
against the CPU oracle (oracle/local_inertial_np.py).

Tolerances: the float32 kernels restate the oracle expression by expression
with FMA contraction off and sum in the same orders, so per-step differences
come only from the last ulp of pow/exp between ROCm's device library and NumPy:
  * one surface step: |dh|, |dhu|, |dhv| <= 2e-6;
  * coupled steps (drainage, every boundary kind, inactive cells): |dh| <= 1e-5 m
    per step, node volume and link flow 1e-4 relative;
  * 1,000 fixed-dt steps on the 20,000-triangle workload: max |dh| <= 1e-4 m;
  * FP64 kernels against the FP64 oracle: |dh| <= 1e-10 m after 100 steps.
"""

import numpy as np
import pytest

from floodhip import synth
from floodhip.solver import HipEnsemble, HipParitySolver
from helpers import basin, oracle_step, small_network_scenario, strip_network, to_oracle
from oracle import local_inertial_np as li
from oracle import v1_parity_np as orc

pytestmark = pytest.mark.gpu
LI = "local_inertial"


def _solver(inp, **kw):
    return HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing,
                           inp.initial_state, surface_mode=LI, **kw)


def _wet_bump_small(n=40):
    inp = synth.make_plain_small(n=n, extent=2000.0)
    h0 = np.maximum(0.0, 2.0 - inp.surface.bed + 0.3 * np.sin(inp.host["centroids"][:, 0] / 150.0))
    hu = (0.2 * np.cos(inp.host["centroids"][:, 1] / 200.0) * h0).astype(np.float32)
    return inp._replace(initial_state=inp.initial_state._replace(h=h0.astype(np.float32), hu=hu))


def test_single_step_matches_oracle():
    inp = _wet_bump_small()
    io = to_oracle(inp)
    rate = 50.0 / 3.6e6
    ref, rdiag = oracle_step(io, io[4], 0.5, rain=rate, surface_mode=LI)
    with _solver(inp) as s:
        diag = s.step(0.5, rate)
        got = s.get_state()
    assert np.max(np.abs(got.h - ref.h)) <= 2e-6
    assert np.max(np.abs(got.hu - ref.hu)) <= 2e-6
    assert np.max(np.abs(got.hv - ref.hv)) <= 2e-6
    assert rdiag["boundary_outflow"] > 0
    for k in ("rainfall_input", "boundary_outflow", "boundary_inflow", "boundary_net_flux", "infiltration_loss"):
        r = rdiag[k]
        assert abs(getattr(diag, k) - r) <= 1e-5 * max(abs(r), 1e-6), (k, getattr(diag, k), r)


def _compare(got, ref, diag=None, rdiag=None):
    assert np.max(np.abs(got.h - ref.h)) <= 1e-5
    assert np.max(np.abs(got.hu - ref.hu)) <= 1e-4
    vs = np.maximum(np.abs(ref.node_volume), 1.0)
    assert np.max(np.abs(got.node_volume - ref.node_volume) / vs, initial=0.0) <= 1e-4
    qs = np.maximum(np.abs(ref.link_flow), 1e-3)
    assert np.max(np.abs(got.link_flow - ref.link_flow) / qs, initial=0.0) <= 1e-4
    if diag is not None:
        for k in orc.LEDGER_FIELDS:
            r = rdiag[k]
            assert abs(getattr(diag, k) - r) <= 1e-4 * max(abs(r), 1e-3), (k, getattr(diag, k), r)


@pytest.mark.parametrize("make", [
    lambda: small_network_scenario(seed=7, h_init=0.3),
    lambda: strip_network(small_network_scenario(seed=8, h_init=0.3), links=True, inlets=False),
])
def test_coupled_steps_match_oracle(make):
    inp = make()
    io = to_oracle(inp)
    st = io[4]
    t = 0.0
    with _solver(inp) as s:
        for i in range(25):
            dt = 0.3 if i % 2 else 0.45
            rate = orc.rain_rate(io[5], t)
            q = orc.external_q(io[5], t + 0.5 * dt)
            st, rdiag = oracle_step(io, st, dt, rain=rate, q=q, surface_mode=LI)
            diag = s.step(dt, rate, q)
            t += dt
            _compare(s.get_state(), st, diag, rdiag)


def test_fp64_kernels_match_fp64_oracle():
    inp = small_network_scenario(seed=21, h_init=0.2, with_external=False)
    io = to_oracle(inp, np.float64)
    st = io[4]
    dt, n = 0.4, 100
    for i in range(n):
        st, _ = oracle_step(io, st, dt, rain=orc.rain_rate(io[5], i * dt), dtype=np.float64, surface_mode=LI)
    with _solver(inp, precision=64, fixed_dt=dt) as s:
        clock, _ = s.advance_to(n * dt - 1e-6, 1e9, max_steps=n)
        h, hu, hv = s.get_state_f64()
    assert clock["global_step"] == n
    assert np.max(np.abs(h - st.h)) <= 1e-10
    assert np.max(np.abs(hu - st.hu)) <= 1e-10


def test_fixed_dt_1000_steps_max_abs_depth_1e4():
    inp = synth.make_plain_small()          # 20,000 triangles, 50 mm/h for 600 s
    dt, steps = 0.5, 1000
    io = to_oracle(inp)
    st = io[4]
    for i in range(steps):
        st, _ = oracle_step(io, st, dt, rain=orc.rain_rate(io[5], i * dt), surface_mode=LI)
    with _solver(inp, fixed_dt=dt) as s:
        clock, _ = s.advance_to(steps * dt, 1.0e9, max_steps=steps)
        got = s.get_state()
    assert clock["global_step"] == steps
    err = np.max(np.abs(got.h - st.h))
    print("local-inertial max |dh| after 1000 fixed-dt steps:", err)
    assert err <= 1e-4


def test_adaptive_dt_policy_matches_oracle():
    inp = _wet_bump_small(n=30)
    io = to_oracle(inp)
    st, clock, led = orc.advance_to_time(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 40.0, 60.0,
                                         *io[:4], io[5], max_steps=10_000, surface_mode=LI)
    with _solver(inp) as s:
        c, ledger = s.advance_to(40.0, 60.0, max_steps=10_000)
        got = s.get_state()
    assert c["global_step"] == clock.global_step
    assert c["cfl_evaluations"] >= 2
    assert abs(c["time_s"] - clock.time_s) < 1e-9 * max(1.0, clock.time_s)
    assert np.max(np.abs(got.h - st.h)) <= 1e-4


def test_stable_timestep_is_the_gravity_wave_timestep():
    inp = _wet_bump_small()
    io = to_oracle(inp)
    ref = li.stable_timestep(io[4], io[0])
    with _solver(inp) as s:
        got = s.stable_timestep()
    assert got == pytest.approx(float(ref), rel=1e-6)
    assert got > orc.stable_timestep(io[4], io[0])     # no velocity term


def test_graph_equals_eager_bitwise():
    inp = small_network_scenario(seed=3, h_init=0.3)
    outs = []
    for use_graph in (True, True, False):
        with _solver(inp, use_graph=use_graph, steps_per_launch=8) as s:
            s.advance_to(30.0, 60.0)
            st = s.get_state()
            outs.append(b"".join(np.asarray(x).tobytes() for x in st) + np.asarray(tuple(s.ledger())).tobytes())
    assert outs[0] == outs[1] == outs[2]


def test_batch_members_equal_plain_runs_bitwise():
    a = _wet_bump_small(n=20)
    b = small_network_scenario(seed=5, h_init=0.3)
    c = strip_network(small_network_scenario(seed=9, h_init=0.3), links=True, inlets=False)
    members = [a, b, c]
    ref = []
    for m in members:
        with _solver(m) as s:
            clock, ledger = s.advance_to(25.0, 600.0, max_steps=100_000)
            ref.append((s.get_state(), clock, ledger))
    with HipEnsemble(members, steps_per_launch=8, surface_mode=LI) as ens:
        ens.advance_to(25.0, 600.0, max_steps=100_000)
        states = ens.member_states()
        clocks = ens.member_clocks()
    for i, (st, (clock, ledger)) in enumerate(zip(states, clocks)):
        rst, rclock, rledger = ref[i]
        assert clock["global_step"] == rclock["global_step"], i
        for name in ("h", "hu", "hv", "wet_time_s", "node_volume", "link_flow", "recharge_storage"):
            assert np.array_equal(getattr(st, name), getattr(rst, name)), (i, name)
        assert np.allclose(np.asarray(ledger), np.asarray(rledger), rtol=1e-12, atol=1e-12), i


def test_lake_at_rest_stays_at_rest():
    bump = lambda c: np.round(4.0 * (1.5 * np.exp(-((c - 60.0) ** 2).sum(1) / 400.0))) / 4.0
    inp = basin(n=10, code="closed", bed=bump, manning=0.03)
    h0 = (np.float32(3.0) - inp.surface.bed).astype(np.float32)
    inp = inp._replace(initial_state=inp.initial_state._replace(h=h0))
    with _solver(inp, fixed_dt=0.5) as s:
        s.advance_to(10.0, 1e9, max_steps=20)
        got = s.get_state()
    assert np.array_equal(got.h, h0)
    assert not np.any(got.hu) and not np.any(got.hv)


def test_compensated_is_closer_to_fp64():
    """The compensated float32 local-inertial step (TwoSum term of h) stays within
    float32 rounding of the plain run and no farther from the FP64 kernels."""
    inp = _wet_bump_small(n=30)
    area = inp.surface.cell_area.astype(np.float64)
    out = {}
    for key, kw in (("f64", dict(precision=64)), ("f32", dict()), ("comp", dict(compensated=True))):
        with _solver(inp, fixed_dt=0.5, **kw) as s:
            s.advance_to(200.0, 1e9, max_steps=400)
            out[key] = s.get_state_f64()[0]
    l2 = lambda a: float(np.sqrt(np.sum((a - out["f64"]) ** 2 * area) / area.sum()))  # noqa: E731
    assert np.max(np.abs(out["comp"] - out["f32"])) <= 1e-4
    assert l2(out["comp"]) <= l2(out["f32"]) * 1.05
