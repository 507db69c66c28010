"""GPU parity of the Dirichlet stage boundary (code 3 with stage values; V1
semantics of gurugram_flood/kernels/triangular_backend.py:1294-1327) against the
CPU oracle, in the coupled step and under the CFL policy.

Tolerances as in the other coupled tests: float32 per step |dh| <= 1e-5 m and
ledger terms within 1e-4 relative; FP64 kernels against the FP64 oracle
<= 1e-10 m.
"""

import numpy as np
import pytest

from floodhip.solver import HipEnsemble, HipParitySolver
from helpers import basin, oracle_step, small_network_scenario, to_oracle
from oracle import v1_parity_np as orc

pytestmark = pytest.mark.gpu


def _stage_scenario(seed=7):
    """The coupled scenario with a third of its boundary edges turned into stage
    edges, each with its own stage: some above the local ground (inflow), some
    below (a dry ghost that lets water out)."""
    inp = small_network_scenario(seed=seed)
    s = inp.surface
    code = s.boundary_code.copy()
    b = np.flatnonzero(s.edge_right < 0)
    stage_edges = b[::3]
    code[stage_edges] = 3
    rng = np.random.default_rng(seed)
    stage = np.zeros(code.shape[0], np.float32)
    z = s.bed[s.edge_left[stage_edges]]
    stage[stage_edges] = (z + rng.uniform(-0.3, 0.4, stage_edges.size)).astype(np.float32)
    return inp._replace(surface=s._replace(boundary_code=code)), stage


@pytest.mark.parametrize("precision", [32, 64])
def test_stage_steps_match_oracle(precision):
    inp, stage = _stage_scenario()
    dtype = np.float32 if precision == 32 else np.float64
    io = to_oracle(inp, dtype)
    st = io[4]
    t = 0.0
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                         boundary_stage_m=stage, precision=precision) as s:
        for i in range(25):
            dt = 0.3 if i % 2 else 0.45
            rate = orc.rain_rate(io[5], t)
            q = orc.external_q(io[5], t + 0.5 * dt)
            st, rdiag = oracle_step(io, st, dt, rain=rate, q=q, dtype=dtype, boundary_stage_m=stage)
            diag = s.step(dt, rate, q)
            t += dt
            if precision == 64:
                h, hu, hv = s.get_state_f64()
                assert np.max(np.abs(h - st.h)) <= 1e-10, i
                assert np.max(np.abs(hu - st.hu)) <= 1e-10, i
            else:
                got = s.get_state()
                assert np.max(np.abs(got.h - st.h)) <= 1e-5, i
            for k in ("boundary_inflow", "boundary_outflow", "rainfall_input", "surface_capture"):
                r = rdiag[k]
                assert abs(getattr(diag, k) - r) <= 1e-4 * max(abs(r), 1e-3), (i, k, getattr(diag, k), r)
    assert rdiag["boundary_inflow"] > 0 and rdiag["boundary_outflow"] > 0


@pytest.mark.parametrize("precision", [32, 64])
def test_stage_edges_of_inactive_cells_match_oracle(precision):
    """Stage edges whose cell is inactive: the stage flows into the inactive cell,
    counts as boundary inflow and is removed with the cell's water; the stage can
    be set and removed after upload."""
    inp, stage = _stage_scenario(seed=5)
    s = inp.surface
    st_edges = np.flatnonzero(s.boundary_code == 3)
    above = st_edges[stage[st_edges] > s.bed[s.edge_left[st_edges]]]
    active = s.active.copy()
    active[s.edge_left[above[::2]]] = 0
    inp = inp._replace(surface=s._replace(active=active))
    dtype = np.float32 if precision == 32 else np.float64
    io = to_oracle(inp, dtype)
    dead = np.flatnonzero((io[0].edge_right < 0) & (io[0].active[io[0].edge_left] < 0.5) & (io[0].boundary_code == 3))
    assert dead.size >= 3
    st = io[4]
    tol_h = 1e-5 if precision == 32 else 1e-10
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                         precision=precision) as sol:
        for i in range(16):
            staged = i >= 4 and i < 12       # no stage, stage values, removed again
            if i == 4:
                sol.set_boundary_stage(stage)
            if i == 12:
                sol.set_boundary_stage(None)
            sm = stage if staged else None
            st, rdiag = oracle_step(io, st, 0.4, rain=2e-5, dtype=dtype, boundary_stage_m=sm)
            diag = sol.step(0.4, 2e-5)
            got = sol.get_state_f64()[0] if precision == 64 else sol.get_state().h
            assert np.max(np.abs(got - st.h)) <= tol_h, i
            for k in ("boundary_inflow", "boundary_outflow", "rainfall_input"):
                r = rdiag[k]
                assert abs(getattr(diag, k) - r) <= 1e-4 * max(abs(r), 1e-3), (i, k, getattr(diag, k), r)
            if staged:
                assert rdiag["boundary_inflow"] > 0


def test_stage_policy_advance_and_scalar_stage():
    inp = basin(n=24, extent=480.0, code="stage", bed=lambda c: 0.002 * c[:, 0],
                h0=lambda c: np.full(c.shape[0], 0.3))
    io = to_oracle(inp)
    stage = 0.9     # above the ground on the low side, below it on the high side
    st, clock, led = orc.advance_to_time(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 60.0, 120.0,
                                         *io[:4], io[5], max_steps=10_000, boundary_stage_m=stage)
    with HipParitySolver(inp.surface, forcing=inp.forcing, state=inp.initial_state,
                         boundary_stage_m=stage) as s:
        c, ledger = s.advance_to(60.0, 120.0, max_steps=10_000)
        got = s.get_state()
    assert c["global_step"] == clock.global_step
    assert np.max(np.abs(got.h - st.h)) <= 1e-4
    for k in ("boundary_inflow", "boundary_outflow"):
        assert abs(getattr(ledger, k) - led[k]) <= 1e-4 * max(abs(led[k]), 1e-3), k
    assert led["boundary_inflow"] > 0 and led["boundary_outflow"] > 0


def test_lake_at_stage_level_stays_at_rest_and_stage_can_be_removed():
    bump = lambda c: 0.8 * np.exp(-((c - 40.0) ** 2).sum(1) / 300.0) + 0.004 * c[:, 0]
    inp = basin(n=10, extent=80.0, code="stage", bed=bump)
    h0 = (np.float32(2.0) - inp.surface.bed).astype(np.float32)
    inp = inp._replace(initial_state=inp.initial_state._replace(h=h0))
    io = to_oracle(inp)
    st = io[4]
    with HipParitySolver(inp.surface, forcing=inp.forcing, state=inp.initial_state, boundary_stage_m=2.0) as s:
        for _ in range(20):
            st, _ = oracle_step(io, st, 0.2, boundary_stage_m=2.0)
            s.step(0.2)
        got = s.get_state()
        assert np.max(np.abs(got.h - st.h)) <= 1e-6
        assert np.max(np.abs(got.h - h0)) <= 1e-4
        # without stage values the stage edges copy the left state (v1_parity.py)
        s.set_boundary_stage(None)
        s.set_state(inp.initial_state._replace(h=np.full(h0.shape, 0.1, np.float32)))
        st = io[4]._replace(h=np.full(h0.shape, 0.1, np.float32))
        for _ in range(10):
            st, rd = oracle_step(io, st, 0.2)
            d = s.step(0.2)
            for k in ("boundary_inflow", "boundary_outflow"):
                assert abs(getattr(d, k) - rd[k]) <= 1e-4 * max(abs(rd[k]), 1e-3), k
        assert np.max(np.abs(s.get_state().h - st.h)) <= 1e-5


def test_stage_batch_members_equal_plain_runs():
    (a, sa), (b, sb) = _stage_scenario(seed=7), _stage_scenario(seed=9)
    with HipEnsemble([a, b]) as ens:
        for m, stage in ((0, sa), (1, sb)):
            ens.select(m).set_boundary_stage(stage)
        ens.select(0)
        ens.advance_to(20.0, 60.0, max_steps=5_000)
        states = ens.member_states()
    for inp, stage, got in ((a, sa, states[0]), (b, sb, states[1])):
        with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                             boundary_stage_m=stage) as s:
            s.advance_to(20.0, 60.0, max_steps=5_000)
            ref = s.get_state()
        np.testing.assert_array_equal(got.h, ref.h)
        np.testing.assert_array_equal(got.node_volume, ref.node_volume)
