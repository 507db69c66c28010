"""GPU parity of the surface step (HIP kernels through the C ABI) against the CPU oracle.

Tolerances: float32 kernels restate the oracle expression by expression with
FMA contraction off, so the only per-step differences are the last-ulp
differences of exp/pow between ROCm's device library and NumPy:
  * one step:  |dh| <= 2e-6 m, |dhu|,|dhv| <= 2e-6 m^2/s;
  * 1,000 fixed-dt steps on the 20,000-triangle workload: max |dh| <= 1e-4 m;
  * ledger terms: relative 1e-5.
"""

import numpy as np
import pytest

from floodhip import synth
from floodhip.solver import HipParitySolver
from floodhip.errors import SimulationError
from helpers import basin, oracle_step, small_network_scenario, to_oracle
from oracle import v1_parity_np as orc

pytestmark = pytest.mark.gpu


def _solver(inp, **kw):
    return HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing,
                           inp.initial_state, **kw)


def _wet_bump_small(n=40):
    inp = synth.make_plain_small(n=n, extent=2000.0)
    h0 = np.maximum(0.0, 2.0 - inp.surface.bed + 0.3 * np.sin(inp.host["centroids"][:, 0] / 150.0))
    return inp._replace(initial_state=inp.initial_state._replace(h=h0.astype(np.float32)))


def test_single_step_matches_oracle():
    inp = _wet_bump_small()
    io = to_oracle(inp)
    rate = 50.0 / 3.6e6
    ref, rdiag = oracle_step(io, io[4], 0.25, rain=rate)
    with _solver(inp) as s:
        diag = s.step(0.25, rate)
        got = s.get_state()
    assert np.max(np.abs(got.h - ref.h)) <= 2e-6
    assert np.max(np.abs(got.hu - ref.hu)) <= 2e-6
    assert np.max(np.abs(got.hv - ref.hv)) <= 2e-6
    for k in ("rainfall_input", "boundary_outflow", "boundary_inflow", "infiltration_loss"):
        r = rdiag[k]
        assert abs(getattr(diag, k) - r) <= 1e-5 * max(abs(r), 1e-6), (k, getattr(diag, k), r)


def test_fixed_dt_1000_steps_max_abs_depth_1e4():
    inp = synth.make_plain_small()          # 20,000 triangles, 50 mm/h for 600 s
    dt, steps = 0.5, 1000
    io = to_oracle(inp)
    st = io[4]
    led = orc.zero_ledger()
    for i in range(steps):
        rate = orc.rain_rate(io[5], i * dt)
        st, d = oracle_step(io, st, dt, rain=rate)
        for k, v in d.items():
            led[k] += v
    with _solver(inp, fixed_dt=dt) as s:
        clock, ledger = s.advance_to(steps * dt, 1.0e9, max_steps=steps)
        got = s.get_state()
    assert clock["global_step"] == steps
    err = np.max(np.abs(got.h - st.h))
    print("max |dh| after 1000 fixed-dt steps:", err)
    assert err <= 1e-4
    assert np.max(np.abs(got.wet_time_s - st.wet_time_s)) <= 1.0
    for k in ("rainfall_input", "boundary_outflow", "boundary_net_flux"):
        r = led[k]
        assert abs(getattr(ledger, k) - r) <= 1e-4 * max(abs(r), 1.0), (k, getattr(ledger, k), r)


def test_adaptive_dt_policy_matches_oracle():
    inp = _wet_bump_small(n=30)
    io = to_oracle(inp)
    st, clock, led = orc.advance_to_time(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 40.0, 60.0,
                                         *io[:4], io[5], max_steps=10_000)
    with _solver(inp) as s:
        c, ledger = s.advance_to(40.0, 60.0, max_steps=10_000)
        got = s.get_state()
    assert c["global_step"] == clock.global_step
    assert abs(c["time_s"] - clock.time_s) < 1e-9 * max(1.0, clock.time_s)
    assert np.max(np.abs(got.h - st.h)) <= 1e-4


def test_bitwise_repeatable_and_graph_equals_eager():
    inp = _wet_bump_small(n=30)
    outs = []
    for use_graph in (True, True, False):
        with _solver(inp, use_graph=use_graph, steps_per_launch=16) as s:
            s.advance_to(30.0, 60.0)
            st = s.get_state()
            outs.append((st.h.tobytes(), st.hu.tobytes(), st.hv.tobytes(), tuple(s.ledger())))
    assert outs[0] == outs[1] == outs[2]


def test_inactive_cells_and_all_boundary_kinds():
    inp = small_network_scenario(h_init=0.4, with_entries=False, with_recharge=False, with_external=False)
    from floodhip.types import empty_coupling, empty_pipe
    inp = inp._replace(coupling=empty_coupling(), pipe=empty_pipe(),
                       initial_state=inp.initial_state._replace(node_volume=np.zeros(0, np.float32),
                                                                link_flow=np.zeros(0, np.float32)))
    io = to_oracle(inp)
    st = io[4]
    with _solver(inp) as s:
        for i in range(20):
            st, d = oracle_step(io, st, 0.2, rain=1e-5)
            diag = s.step(0.2, 1e-5)
        got = s.get_state()
    assert np.max(np.abs(got.h - st.h)) <= 1e-5
    assert np.all(got.h[~inp.surface.active] == 0)


def test_stable_timestep_matches_oracle():
    inp = _wet_bump_small()
    io = to_oracle(inp)
    ref = orc.stable_timestep(io[4], io[0])
    with _solver(inp) as s:
        got = s.stable_timestep()
        c = s.clock()
    assert got == pytest.approx(float(ref), rel=1e-6)
    assert c["global_step"] == 0 and c["time_s"] == 0.0


def test_failure_modes_raise_reference_errors():
    inp = _wet_bump_small(n=10)
    with _solver(inp) as s:
        with pytest.raises(ValueError):
            s.step(0.0)
        with pytest.raises(SimulationError):
            s.advance_to(100.0, 100.0, max_steps=3)
        h = inp.initial_state.h.copy()
        h[5] = np.nan
        s.set_state(inp.initial_state._replace(h=h))
        with pytest.raises(SimulationError):
            s.advance_to(s.clock()["time_s"] + 1.0, 1000.0)


def test_mode_and_dry_weather_argument_errors():
    """Unknown surface modes, a dry-weather step on a batch, a non-positive dt and a
    non-finite tailwater override raise ValueError; the context stays usable."""
    from floodhip.solver import HipEnsemble
    inp = small_network_scenario(seed=4, h_init=0.2)
    with pytest.raises(ValueError):
        _solver(inp, surface_mode="diffusive")
    with _solver(inp) as s:
        n = inp.coupling.node_storage_area.shape[0]
        with pytest.raises(ValueError):
            s.dry_weather_step(0.0)
        with pytest.raises(ValueError):
            s.dry_weather_step(1.0, outfall_tailwater_head_m=np.full(n, np.nan, np.float32))
        d = s.dry_weather_step(1.0, np.full(n, 0.01, np.float32))
        assert d.dry_weather_inflow_m3 == pytest.approx(0.01 * n, rel=1e-6)
    with HipEnsemble([inp, inp]) as ens:
        with pytest.raises(ValueError):
            ens.dry_weather_step(1.0)


@pytest.mark.parametrize("mode", ["full_swe", "local_inertial"])
def test_cfl_minimum_does_not_leak_between_calls(mode):
    """The step's CFL minimum lives in per-block slots.  A stable_timestep() call on a
    fast, deep state followed by a calm state must leave the advance exactly as a
    solver that only ever held the calm state: same steps, dt, time and depths."""
    calm = _wet_bump_small()
    deep = calm.initial_state._replace(h=(calm.initial_state.h + np.float32(8.0)).astype(np.float32),
                                       hu=np.full_like(calm.initial_state.hu, 3.0))
    with _solver(calm, surface_mode=mode) as ref, _solver(calm, surface_mode=mode) as s:
        ref_dt0 = ref.stable_timestep()
        s.set_state(deep)
        fast_dt = s.stable_timestep()
        assert fast_dt < 0.5 * ref_dt0          # the deep state's minimum is much smaller
        s.set_state(calm.initial_state)
        for target in (20.0, 45.0):
            ref.advance_to(target, 600.0)
            s.advance_to(target, 600.0)
            assert s.clock() == ref.clock()
            assert np.array_equal(s.get_state().h, ref.get_state().h)
        assert s.stable_timestep() == ref.stable_timestep()
