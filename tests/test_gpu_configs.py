"""The benchmark workloads at their full sizes on the GPU.

* 981,880-triangle coupled production workload: 20 fixed 0.5 s steps, every step
  against the CPU oracle (depth, node volume, link flow, ledger terms); the
  one-step error along 120 steps of the oracle's CFL trajectory; the FP64
  ledger's balance identity (parity_runner.py:337-345); finite, non-negative
  depths; bit-identical repeat runs, graph and eager.
* 200,000-triangle dam break, 120 s, in FP32, compensated FP32 and FP64 against
  each other and Stoker's solution.
* 48 uneven catchments (20k-300k triangles, 5.4 M in all) in one batch: every
  member bit-identical to its own plain run.
* 16 members on the production mesh in one compensated, graph-replayed batch:
  every member bit-identical to its own plain run; stored volume within 1e-5
  relative of the same batch in FP64.

Tolerances of the per-step comparison (float32, device-library transcendentals
against NumPy's): |dh| <= 1e-5 m; node volume within 1e-4 of max(V, 1 m^3);
link flow within 1e-3 of max(|Q|, 0.01 m^3/s); ledger terms within 1e-4 relative.
"""

import json

import numpy as np
import pytest

from floodhip import synth
from floodhip.precision import ensemble_vs_fp64, study
from floodhip.solver import HipEnsemble, HipParitySolver
from helpers import oracle_step, to_oracle
from oracle import v1_parity_np as orc

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]


def _report(name, **kw):
    print(f"\n[{name}] " + json.dumps(kw, default=float), flush=True)


def _solver(inp, **kw):
    return HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state, **kw)


@pytest.fixture(scope="module")
def production():
    return synth.make_production_scale()


def _balance_residual(inp, st, ledger):
    """parity_runner.py:337-345 over one advance from the initial state (no recharge rows)."""
    area = inp.surface.cell_area.astype(np.float64)
    v_s0 = float(np.sum(inp.initial_state.h.astype(np.float64) * area))
    v_n0 = float(np.sum(inp.initial_state.node_volume.astype(np.float64)))
    lhs = v_s0 + v_n0 + ledger.rainfall_input + ledger.external_boundary_inflow
    rhs = (float(np.sum(st.h.astype(np.float64) * area)) + float(np.sum(st.node_volume.astype(np.float64)))
           + ledger.infiltration_loss + ledger.outfall_discharge + ledger.unreturned_overflow
           + ledger.recharge_loss - ledger.boundary_net_flux)
    return (lhs - rhs) / max(abs(lhs), 1.0)


def test_production_fixed_steps_match_oracle(production):
    inp = production
    io = to_oracle(inp)
    st = io[4]
    dt, t = 0.5, 0.0
    worst = {"dh": 0.0, "node": 0.0, "link": 0.0, "ledger": 0.0}
    with _solver(inp) as s:
        for i in range(20):
            rate = orc.rain_rate(io[5], t)
            q = orc.external_q(io[5], t + 0.5 * dt)
            st, rdiag = oracle_step(io, st, dt, rain=rate, q=q)
            diag = s.step(dt, rate, q)
            t += dt
            got = s.get_state()
            dh = float(np.max(np.abs(got.h - st.h)))
            node = float(np.max(np.abs(got.node_volume - st.node_volume) / np.maximum(np.abs(st.node_volume), 1.0)))
            link = float(np.max(np.abs(got.link_flow - st.link_flow) / np.maximum(np.abs(st.link_flow), 1e-2)))
            led = max(abs(getattr(diag, k) - rdiag[k]) / max(abs(rdiag[k]), 1e-3) for k in orc.LEDGER_FIELDS)
            for k, v in (("dh", dh), ("node", node), ("link", link), ("ledger", led)):
                worst[k] = max(worst[k], v)
            assert dh <= 1e-5 and node <= 1e-4 and link <= 1e-3 and led <= 1e-4, (i, dh, node, link, led)
        ledger = s.ledger()
        got = s.get_state()
    assert np.isfinite(got.h).all() and got.h.min() >= -1e-7
    residual = _balance_residual(inp, got, ledger)
    _report("config2_fixed_steps", steps=20, **worst, balance_residual_ratio=residual,
            rain_m3=ledger.rainfall_input, capture_m3=ledger.surface_capture)
    assert abs(residual) <= 1e-6


def test_production_one_step_error_along_the_oracle_trajectory(production):
    """Along 120 steps of the oracle's CFL trajectory (the C restatement, all host
    cores) the GPU restarts from the oracle's float32 state before every step and
    takes the same step: the one-step error stays at transcendental rounding, at
    outfall nodes too (free-running trajectories part at the outfall threshold)."""
    from oracle import v1_parity_c as oc
    inp = production
    s_, c_, p_, r_, st, f_ = to_oracle(inp)
    co = oc.COracle(s_, c_, p_, r_, f_)
    outfall = np.asarray(inp.pipe.outfall_area) > 0
    clock, ledger = orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger()
    worst = {"dh": 0.0, "node": 0.0, "outfall_node": 0.0, "link": 0.0}
    with _solver(inp) as s:
        for _ in range(120):
            before = st
            st, c1, ledger = co.advance_to_time(st, clock, ledger, 3600.0, 3600.0, max_steps=1)
            s.set_state(before)
            s.step(c1.time_s - clock.time_s, orc.rain_rate(f_, clock.time_s))
            got = s.get_state()
            rel = np.abs(got.node_volume.astype(np.float64) - st.node_volume) / np.maximum(np.abs(st.node_volume), 1.0)
            q = np.abs(got.link_flow.astype(np.float64) - st.link_flow) / np.maximum(np.abs(st.link_flow), 1e-2)
            for k, v in (("dh", float(np.max(np.abs(got.h - st.h)))), ("node", float(rel.max())),
                         ("outfall_node", float(rel[outfall].max())), ("link", float(q.max()))):
                worst[k] = max(worst[k], v)
            clock = c1
    _report("config2_one_step_error", steps=120, t=clock.time_s, **worst)
    assert worst["dh"] <= 1e-6 and worst["node"] <= 1e-5 and worst["outfall_node"] <= 1e-5 and worst["link"] <= 1e-3


def test_production_policy_run_is_repeatable_and_closes_in_fp64(production):
    inp = production
    runs = []
    for graph in (True, True, False):
        with _solver(inp, use_graph=graph) as s:
            c, ledger = s.advance_to(3600.0, 3600.0, max_steps=200, raise_on_budget=False)
            runs.append((s.get_state(), c, ledger))
    (a, ca, la), (b, cb, lb), (e, ce, le) = runs
    assert ca["global_step"] == cb["global_step"] == ce["global_step"] == 200
    for x in (b, e):
        np.testing.assert_array_equal(a.h, x.h)
        np.testing.assert_array_equal(a.hu, x.hu)
        np.testing.assert_array_equal(a.node_volume, x.node_volume)
        np.testing.assert_array_equal(a.link_flow, x.link_flow)
    assert tuple(la) == tuple(lb) == tuple(le)
    assert np.isfinite(a.h).all() and a.h.min() >= -1e-7
    r32 = _balance_residual(inp, a, la)
    with _solver(inp, precision=64) as s:
        s.advance_to(3600.0, 3600.0, max_steps=200, raise_on_budget=False)
        h64 = s.get_state_f64()[0]
        st64, l64 = s.get_state(), s.ledger()
    area = inp.surface.cell_area.astype(np.float64)
    v_s0 = float(np.sum(inp.initial_state.h.astype(np.float64) * area))
    v_n0 = float(np.sum(inp.initial_state.node_volume.astype(np.float64)))
    lhs = v_s0 + v_n0 + l64.rainfall_input + l64.external_boundary_inflow
    rhs = (float(np.sum(h64 * area)) + float(np.sum(st64.node_volume.astype(np.float64))) + l64.infiltration_loss
           + l64.outfall_discharge + l64.unreturned_overflow + l64.recharge_loss - l64.boundary_net_flux)
    # node volumes come back as float32 copies of the FP64 state: 1e-7 of the stored total
    r64 = (lhs - rhs) / max(abs(lhs), 1.0)
    stored = v_n0 + float(np.sum(h64 * area))
    _report("config2_policy", steps=200, t=ca["time_s"], balance_residual_f32=r32, balance_residual_f64=r64,
            depth_linf_vs_fp64=float(np.max(np.abs(a.h - h64))))
    assert abs(r32) <= 1e-6
    assert abs(lhs - rhs) <= 1e-7 * stored


def test_dam_break_precision_study_200k():
    res = study(120.0)
    f = res["forms"]
    _report("config3_precision", **{k: {kk: v[kk] for kk in ("steps", "depth_l2_vs_fp64_m", "depth_linf_vs_fp64_m",
                                                             "depth_l1_vs_stoker_m", "relative_volume_drift")}
                                    for k, v in f.items()})
    assert f["fp64"]["steps"] > 1000
    for name in f:
        assert f[name]["depth_l1_vs_stoker_m"] < 0.35, name      # numerical diffusion at the front
    assert abs(f["fp64"]["relative_volume_drift"]) <= 1e-12
    assert abs(f["fp32_compensated"]["relative_volume_drift"]) <= 1e-6
    assert f["fp32_compensated"]["depth_l2_vs_fp64_m"] <= f["fp32_plain"]["depth_l2_vs_fp64_m"]
    assert f["fp32_plain"]["depth_linf_vs_fp64_m"] <= 1e-2


def _plain_states(members, steps, **opts):
    out = []
    for inp in members:
        with _solver(inp, **opts) as s:
            s.advance_to(3600.0, 3600.0, max_steps=steps, raise_on_budget=False)
            out.append((s.get_state(), s.clock()["global_step"], s.ledger()))
    return out


def _check_members(batch_states, batch_clocks, plain):
    for m, (got, (c, led)) in enumerate(zip(batch_states, batch_clocks)):
        ref, steps, rled = plain[m]
        assert c["global_step"] == steps, m
        np.testing.assert_array_equal(got.h, ref.h, err_msg=f"member {m}")
        np.testing.assert_array_equal(got.hu, ref.hu, err_msg=f"member {m}")
        np.testing.assert_array_equal(got.node_volume, ref.node_volume, err_msg=f"member {m}")
        np.testing.assert_array_equal(got.link_flow, ref.link_flow, err_msg=f"member {m}")
        assert tuple(led) == tuple(rled), m


def test_uneven_batch_of_48_members_equals_plain_runs():
    members = synth.make_uneven_ensemble()
    assert len(members) == 48
    steps = 40
    opts = dict(compensated=True, pack_order=2, block_size=128)
    with HipEnsemble(members, **opts) as ens:
        ens.advance_to(3600.0, 3600.0, max_steps=steps, raise_on_budget=False)
        states, clocks = ens.member_states(), ens.member_clocks()
    # the same launch shape (block size) as the batch, so the ledger partials line up
    plain = _plain_states(members, steps, compensated=True, block_size=128)
    _check_members(states, clocks, plain)
    _report("config4_uneven", members=48, triangles=int(sum(m.surface.cell_area.shape[0] for m in members)),
            steps=steps, min_dt_steps=min(c["global_step"] for c, _ in clocks))


def test_shared_mesh_batch_of_16_members_equals_plain_runs_and_fp64(production):
    members = synth.make_shared_mesh_ensemble(production, n_members=16)
    steps = 48
    # one launch shape for the batch and the plain runs, so the ledger partials line up
    opts = dict(compensated=True, use_graph=True, edges_per_thread=2, block_size=256)
    with HipEnsemble(members, **opts) as ens:
        ens.advance_to(3600.0, 3600.0, max_steps=steps, raise_on_budget=False)
        states, clocks = ens.member_states(), ens.member_clocks()
    plain = _plain_states(members, steps, **opts)
    _check_members(states, clocks, plain)
    fp64 = ensemble_vs_fp64(members, steps=120, dt=0.5, use_graph=True)
    _report("config5_shared", members=16, steps=steps, vs_fp64=fp64["worst"])
    assert fp64["worst"]["stored_volume_rel"] <= 1e-5
    assert fp64["worst"]["depth_linf_m"] <= 1e-4
