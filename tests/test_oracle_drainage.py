"""Pins the CPU oracle's drainage processes (v1_parity.py:259-427) against
geometry closed forms and per-process conservation identities."""

import numpy as np
import pytest

from helpers import small_network_scenario, to_oracle
from oracle import v1_parity_np as orc

F64 = np.float64


@pytest.mark.parametrize("D", [0.3, 1.0, 2.0])
def test_circular_section_closed_forms(D):
    d = np.array([D], F64)
    assert np.isclose(orc._circular_area(0.5 * d, d, F64)[0], np.pi * D * D / 8, rtol=1e-14)
    assert np.isclose(orc._circular_area(d, d, F64)[0], np.pi * D * D / 4, rtol=1e-14)
    assert np.isclose(orc._circular_area(2 * d, d, F64)[0], np.pi * D * D / 4, rtol=1e-14)
    assert np.isclose(orc._circular_radius(0.5 * d, d, F64)[0], D / 4, rtol=1e-14)
    assert np.isclose(orc._circular_radius(2 * d, d, F64)[0], D / 4, rtol=1e-14)
    assert orc._circular_area(np.zeros(1), d, F64)[0] == 0.0
    # segment area for a quarter depth: r^2 acos((r-h)/r) - (r-h) sqrt(2rh - h^2)
    r, h = D / 2, D / 4
    seg = r * r * np.arccos((r - h) / r) - (r - h) * np.sqrt(2 * r * h - h * h)
    assert np.isclose(orc._circular_area(np.array([h]), d, F64)[0], seg, rtol=1e-12)


def _scenario(dtype=F64, **kw):
    inp = small_network_scenario(**kw)
    return inp, to_oracle(inp, dtype)


def test_inlet_exchange_conserves_and_limits_donors():
    inp, (s, c, p, r, st, f) = _scenario()
    A = s.cell_area
    # deep ponding on inlet cells, nearly empty nodes, and a huge dt so requests exceed what is available
    h = st.h.copy()
    h[c.inlet_cell] = 0.02
    st = st._replace(h=h, node_volume=st.node_volume * 0.01)
    new, cap, sur = orc.inlet_exchange(st, s, c, 500.0, F64)
    assert cap > 0 and sur >= 0
    dvs = np.sum((new.h - st.h) * A)
    dvn = np.sum(new.node_volume - st.node_volume)
    assert np.isclose(dvs, sur - cap, rtol=1e-12, atol=1e-12)
    assert np.isclose(dvn, cap - sur, rtol=1e-12, atol=1e-12)
    assert np.all(new.h >= 0) and np.all(new.node_volume >= 0)
    # a cell cannot give more than it holds
    for cell in np.unique(c.inlet_cell):
        assert new.h[cell] >= -1e-15


def test_links_conserve_node_volume_and_respect_donors():
    inp, (s, c, p, r, st, f) = _scenario()
    # no outfalls: transfers only move volume between nodes
    p = p._replace(outfall_area=np.zeros_like(p.outfall_area))
    q0 = np.random.default_rng(1).normal(0.0, 2.0, p.link_from.shape[0])
    st = st._replace(link_flow=q0)
    new, outfall = orc.pipe_links_and_outfalls(st, c, p, 50.0, F64)
    assert outfall == 0.0
    assert np.isclose(new.node_volume.sum(), st.node_volume.sum(), rtol=1e-12)
    assert np.all(new.node_volume >= 0)
    # every link flow is within capacity
    assert np.all(np.abs(new.link_flow) <= np.maximum(p.link_capacity, 0) * (1 + 1e-12))


def test_outfall_removes_exactly_the_reported_volume():
    inp, (s, c, p, r, st, f) = _scenario()
    new, outfall = orc.pipe_links_and_outfalls(st, c, p, 2.0, F64)
    assert outfall > 0
    assert np.isclose(st.node_volume.sum() - new.node_volume.sum(), outfall, rtol=1e-12)


def test_overflow_return_weighted_table_closed_form():
    inp, (s, c, p, r, st, f) = _scenario()
    N = c.node_storage_area.shape[0]
    cap = c.node_storage_area * np.maximum(c.node_max_depth, 0)
    vol = cap.copy()
    vol[0] += 10.0                            # node 0 overflows by 10 m^3 into two weighted cells
    cells = np.array([5, 9], np.int32)
    c = c._replace(surcharge_cell=cells, surcharge_node=np.array([0, 0], np.int32),
                   surcharge_weight=np.array([1.0, 3.0]), surcharge_weight_normalized=False,
                   node_surface_cell=np.full(N, -1, np.int32))
    st = st._replace(node_volume=vol, h=np.zeros_like(st.h))
    new, total, unret = orc.return_node_overflow(st, s, c, F64)
    assert np.isclose(total, 10.0) and unret == 0.0
    assert np.isclose(new.h[5] * s.cell_area[5], 2.5) and np.isclose(new.h[9] * s.cell_area[9], 7.5)
    assert np.isclose(new.node_volume[0], cap[0])


def test_overflow_fallback_and_unreturned():
    inp, (s, c, p, r, st, f) = _scenario(with_entries=False)
    cap = c.node_storage_area * np.maximum(c.node_max_depth, 0)
    vol = cap + 1.0
    st = st._replace(node_volume=vol)
    new, total, unret = orc.return_node_overflow(st, s, c, F64)
    no_cell = c.node_surface_cell < 0
    assert np.isclose(unret, float(no_cell.sum()))
    assert np.isclose(total, float((~no_cell).sum()))
    assert np.isclose(np.sum((new.h - st.h) * s.cell_area), total, rtol=1e-12)


def test_recharge_capture_and_loss_balance():
    inp, (s, c, p, r, st, f) = _scenario()
    st = st._replace(h=np.full_like(st.h, 0.05))
    new, capt, loss = orc.apply_recharge(st, s, r, 10.0, F64)
    assert capt > 0 and loss > 0
    assert np.isclose(np.sum((st.h - new.h) * s.cell_area), capt, rtol=1e-12)
    assert np.isclose(new.recharge_storage.sum() - st.recharge_storage.sum(), capt - loss, rtol=1e-12)
    assert np.all(new.recharge_storage <= r.storage_capacity + 1e-12)


def test_coupled_ledger_balance_identity_f64():
    """parity_runner.py:337-345 balance over a full coupled run (every node has a
    surface cell, so nothing is 'unreturned', and every cell is active: water the
    reference sends onto an inactive cell is zeroed by the next surface step)."""
    inp, (s, c, p, r, st, f) = _scenario(h_init=0.3)
    s = s._replace(active=np.ones_like(s.active))
    c = c._replace(node_surface_cell=np.abs(c.node_surface_cell))
    A = s.cell_area
    v_s0, v_n0, v_r0 = np.sum(st.h * A), st.node_volume.sum(), st.recharge_storage.sum()
    led = orc.zero_ledger()
    t = 0.0
    for i in range(60):
        dt = 0.2
        q = orc.external_q(f, t + 0.5 * dt).astype(F64)
        st, d = orc.advance_parity_step(
            st, s, c, p, r, rainfall_rate_mps=orc.rain_rate(f, t), rainfall_multiplier=f.rainfall_multiplier,
            horton_f0_mps=f.horton_f0_mps, horton_fc_mps=f.horton_fc_mps, horton_k_per_s=f.horton_k_per_s,
            external_receiver_cell=f.external_receiver_cell, external_q_m3ps=q, dt=dt, dtype=F64)
        for k, v in d.items():
            led[k] += v
        t += dt
    lhs = v_s0 + v_n0 + v_r0 + led["rainfall_input"] + led["external_boundary_inflow"]
    rhs = (np.sum(st.h * A) + st.node_volume.sum() + st.recharge_storage.sum() + led["infiltration_loss"]
           + led["outfall_discharge"] + led["recharge_loss"] - led["boundary_net_flux"])
    assert led["unreturned_overflow"] == 0.0
    assert led["rainfall_input"] > 0 and led["external_boundary_inflow"] > 0 and led["outfall_discharge"] > 0
    assert led["surface_capture"] > 0 and led["recharge_capture"] > 0
    assert abs(lhs - rhs) / max(abs(lhs), 1.0) < 1e-13


def test_rain_and_external_lookup():
    inp, (s, c, p, r, st, f) = _scenario()
    assert orc.rain_rate(f, 0.0) == pytest.approx(30.0 / 3.6e6, rel=1e-6)
    assert orc.rain_rate(f, 60.0) == pytest.approx(80.0 / 3.6e6, rel=1e-6)   # [start, end)
    assert orc.rain_rate(f, 179.999) == pytest.approx(10.0 / 3.6e6, rel=1e-6)
    assert orc.rain_rate(f, 180.0) == 0.0
    q = f.external_q_m3ps
    assert np.allclose(orc.external_q(f, 15.0), q[0] + 0.5 * (q[1] - q[0]), rtol=1e-6)
    assert np.allclose(orc.external_q(f, -5.0), q[0])
    assert np.allclose(orc.external_q(f, 1e6), q[-1])


def _network64(seed=7):
    from helpers import small_network_scenario, to_oracle
    inp = small_network_scenario(seed=seed, h_init=0.2)
    s, c, p, r, st, f = to_oracle(inp, F64)
    return c, p, st


def test_dry_weather_balance_with_outfalls_and_cap():
    """triangular_backend.py:833-952: V_after = V_before + inflow - outfall - capped
    overflow, every node at or below its capacity, diagnostics consistent."""
    c, p, st = _network64()
    rng = np.random.default_rng(1)
    inflow = rng.uniform(-0.01, 0.05, c.node_storage_area.shape[0])
    v0 = float(np.sum(np.maximum(st.node_volume, 0)))
    new, d = orc.dry_weather_step(st, c, p, 30.0, inflow, dtype=F64)
    assert d["dry_weather_inflow"] == pytest.approx(float(np.sum(np.maximum(inflow, 0) * 30.0)), rel=1e-14)
    assert d["outfall_discharge"] > 0 and d["link_abs_flow"] > 0 and d["overflow_surcharge"] > 0
    balance = v0 + d["dry_weather_inflow"] - d["outfall_discharge"] - d["overflow_surcharge"]
    assert float(np.sum(new.node_volume)) == pytest.approx(balance, rel=1e-12)
    assert d["final_node_volume"] == pytest.approx(float(np.sum(new.node_volume)), rel=1e-14)
    cap = c.node_storage_area * np.maximum(c.node_max_depth, 0)
    assert np.all(new.node_volume <= cap)


def test_dry_weather_closed_outfalls_conserve_and_uncapped_overflow_stays():
    c, p, st = _network64()
    p = p._replace(outfall_area=np.zeros_like(p.outfall_area))
    inflow = np.full(c.node_storage_area.shape[0], 0.02)
    v0 = float(np.sum(np.maximum(st.node_volume, 0)))
    new, d = orc.dry_weather_step(st, c, p, 60.0, inflow, cap_node_overflow=False, dtype=F64)
    assert d["outfall_discharge"] == 0.0
    assert float(np.sum(new.node_volume)) == pytest.approx(v0 + d["dry_weather_inflow"], rel=1e-12)
    cap = c.node_storage_area * np.maximum(c.node_max_depth, 0)
    assert d["overflow_surcharge"] == pytest.approx(float(np.sum(np.maximum(new.node_volume - cap, 0))), rel=1e-12)
    assert np.any(new.node_volume > cap)


def test_dry_weather_tailwater_above_heads_stops_outfalls():
    c, p, st = _network64()
    high = np.full(c.node_storage_area.shape[0], 1.0e4)
    _, d = orc.dry_weather_step(st, c, p, 30.0, None, outfall_tailwater_head_m=high, dtype=F64)
    assert d["outfall_discharge"] == 0.0 and d["dry_weather_inflow"] == 0.0
