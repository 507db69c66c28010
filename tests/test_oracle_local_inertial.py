"""Pins the CPU oracle's local-inertial surface step (triangular_backend.py:1181-1291,
local_inertial.py:20-48, network_kernels.py:45-82, equations.py:181-207) against
closed forms, conservation laws, donor-limiting invariants and symmetry."""

import numpy as np
import pytest

from helpers import basin, oracle_step, small_network_scenario, to_oracle
from oracle import local_inertial_np as li
from oracle import v1_parity_np as orc

G = orc.GRAVITY


def run(inp, steps, dt, dtype=np.float64, rain=0.0):
    io = to_oracle(inp, dtype)
    st = io[4]
    led = orc.zero_ledger()
    for _ in range(steps):
        st, d = oracle_step(io, st, dt, rain=rain, dtype=dtype, surface_mode="local_inertial")
        for k, v in d.items():
            led[k] += v
    return st, led, io


def _unit_square(code="closed", n=0.0, bed=(0.0, 0.0)):
    from floodhip.topology import build_surface_topology
    nodes = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    tri = np.array([[0, 1, 2], [0, 2, 3]])
    surface, _ = build_surface_topology(nodes, tri, np.asarray(bed, np.float32), n, boundary_mode=code)
    return orc.cast_tuple(orc.SurfaceTopology(*surface), np.float64)


def _state(h, hu=None, hv=None):
    h = np.asarray(h, np.float64)
    z = np.zeros_like(h)
    return orc.ParityState(h, z if hu is None else np.asarray(hu, np.float64),
                           z if hv is None else np.asarray(hv, np.float64), z, np.zeros(0), np.zeros(0), np.zeros(0))


def test_lake_at_rest_is_exact():
    # bed in quarter metres and eta = 3 m: bed + h is exact, so every slope is exactly 0
    bump = lambda c: np.round(4.0 * (1.5 * np.exp(-((c - 60.0) ** 2).sum(1) / 400.0))) / 4.0
    inp = basin(n=10, code="closed", bed=bump, manning=0.03)
    assert len(np.unique(inp.surface.bed)) > 3
    for dtype in (np.float64, np.float32):
        io = list(to_oracle(inp, dtype))
        h0 = (dtype(3.0) - io[0].bed).astype(dtype)
        io[4] = io[4]._replace(h=h0)
        st = io[4]
        for _ in range(20):
            st, d = oracle_step(io, st, 0.5, dtype=dtype, surface_mode="local_inertial")
        assert np.array_equal(st.h, h0)
        assert not np.any(st.hu) and not np.any(st.hv)


def test_closed_basin_conserves_volume():
    hump = lambda c: 0.5 + 0.4 * np.exp(-((c - 40.0) ** 2).sum(1) / 300.0)
    inp = basin(n=12, code="closed", h0=hump, manning=0.02)
    area = inp.surface.cell_area.astype(np.float64)
    v0 = np.sum(inp.initial_state.h.astype(np.float64) * area)
    st64, led, _ = run(inp, 60, 0.2, np.float64)
    assert abs(np.sum(st64.h * area) - v0) / v0 < 1e-12
    assert led["boundary_outflow"] == 0.0
    st32, _, _ = run(inp, 60, 0.2, np.float32)
    assert abs(np.sum(st32.h.astype(np.float64) * area) - v0) / v0 < 2e-6
    assert np.max(np.abs(st32.h - st64.h)) < 1e-4
    assert np.max(np.abs(st64.hu)) > 1e-3     # the hump does spread


def test_two_cell_step_by_hand():
    """Flat bed, at rest, h 1.0 | 0.5 across the unit square's diagonal: Q_old = 0, so
    Q = -g (h_flow L) dt (eta_R - eta_L) / d with h_flow = 1, d = mean Lc; the left
    cell donates Q dt (well below its volume); momentum is Q n / (2 A / Lc)."""
    s = _unit_square(n=0.03)
    dt = 0.01
    new, bnet, bin_, bout = li.surface_step(_state([1.0, 0.5]), s, dt, np.float64)
    e = int(np.flatnonzero(s.edge_right >= 0)[0])
    left, right = int(s.edge_left[e]), int(s.edge_right[e])
    h = np.array([1.0, 0.5])
    L = np.sqrt(2.0)
    d = 0.5 * (s.characteristic_length[left] + s.characteristic_length[right])
    Q = -G * (1.0 * L) * dt * (h[right] - h[left]) / d
    A = 0.5
    assert np.isclose(new.h[left], h[left] - Q * dt / A, rtol=1e-13)
    assert np.isclose(new.h[right], h[right] + Q * dt / A, rtol=1e-13)
    nx, ny = s.edge_nx[e], s.edge_ny[e]
    for c in (left, right):
        per = 2.0 * A / s.characteristic_length[c]
        assert np.isclose(new.hu[c], Q * nx / per, rtol=1e-13)
        assert np.isclose(new.hv[c], Q * ny / per, rtol=1e-13)
    assert bnet == 0.0 and bin_ == 0.0 and bout == 0.0


def test_friction_closed_form():
    """Slope 0 and Q_old != 0: Q = Q_old / (1 + g dt n^2 |Q_old| / (R^(4/3) A))."""
    q_old = np.array([0.8, -0.3])
    hf, L, n, dt = 0.4, 2.0, 0.05, 0.5
    A = hf * L
    q = li.discharge_update(q_old, np.full(2, A), np.full(2, hf), np.zeros(2), np.full(2, n), dt, np.float64)
    expect = q_old / (1.0 + G * dt * n * n * np.abs(q_old) / (hf ** (4.0 / 3.0) * A))
    assert np.allclose(q, expect, rtol=1e-14)
    # frictionless, still: Q = -g A dt S
    q = li.discharge_update(np.zeros(1), np.full(1, A), np.full(1, hf), np.full(1, 0.01), np.zeros(1), dt,
                            np.float64)
    assert q[0] == pytest.approx(-G * A * dt * 0.01, rel=1e-15)


def test_donor_limiting_drains_at_most_the_volume():
    """A tall column in a closed basin with a huge dt: every requested transfer out of
    it is scaled so its depth ends at 0 (not below), and volume is conserved."""
    col = lambda c: np.where(((c - 60.0) ** 2).sum(1) < 150.0, 5.0, 0.0)
    inp = basin(n=10, code="closed", h0=col, manning=0.0)
    io = to_oracle(inp, np.float64)
    area = io[0].cell_area
    v0 = np.sum(io[4].h * area)
    new, *_ = li.surface_step(io[4], io[0], 50.0, np.float64)
    assert new.h.min() >= 0.0
    assert abs(np.sum(new.h * area) - v0) / v0 < 1e-12
    # the column's rim cells give up their whole volume; its inner cells keep theirs
    tall = io[4].h > 0
    assert np.sum(new.h[tall] < 1e-12) > 0 and np.sum(new.h[tall] == 5.0) > 0


@pytest.mark.parametrize("code,drains", [("outflow", True), ("transmissive", True), ("closed", False),
                                         ("reflective", False)])
def test_boundary_kinds(code, drains):
    """Codes 4 and 1 let water out (never in); 0 and 2 are walls.  The outflow of a
    step never exceeds the volume left after the interior transfers."""
    inp = basin(n=6, code=code, h0=lambda c: np.full(c.shape[0], 1.0), manning=0.0)
    io = to_oracle(inp, np.float64)
    area = io[0].cell_area
    v0 = np.sum(io[4].h * area)
    new, bnet, bin_, bout = li.surface_step(io[4], io[0], 100.0, np.float64)
    assert bin_ == 0.0
    if drains:
        assert bout > 0.0 and bnet == -bout
        assert new.h.min() >= 0.0
        assert np.sum(new.h * area) == pytest.approx(v0 - bout, rel=1e-12)
    else:
        assert bout == 0.0
        assert np.sum(new.h * area) == pytest.approx(v0, rel=1e-12)


def test_mirror_symmetry_about_diagonal():
    hump = lambda c: 0.3 + 0.6 * np.exp(-((c[:, 0] - 30.0) ** 2 + (c[:, 1] - 30.0) ** 2) / 200.0)
    inp = basin(n=10, code="closed", h0=hump, manning=0.01)
    st, _, _ = run(inp, 30, 0.2, np.float64)
    cen = inp.host["centroids"]
    key = {tuple(np.round(c, 6)): i for i, c in enumerate(cen)}
    mirror = np.array([key[tuple(np.round(c[::-1], 6))] for c in cen])
    assert np.allclose(st.h, st.h[mirror], atol=1e-12)
    assert np.allclose(st.hu, st.hv[mirror], atol=1e-12)


def test_stable_timestep_ignores_velocity():
    g = orc.GRAVITY
    lc = np.array([4.0, 10.0, 2.5])
    fields = {f: np.zeros(0) for f in orc.SurfaceTopology._fields}
    fields.update(characteristic_length=lc, cell_area=np.ones(3))
    topo = orc.SurfaceTopology(**fields)
    st = _state(np.array([2.0, 2.0, 0.0]), hu=np.array([0.0, 60.0, 0.0]))
    # a dry cell counts at the 1e-5 m floor
    expect = 0.85 * min(4.0 / np.sqrt(g * 2.0), 10.0 / np.sqrt(g * 2.0), 2.5 / np.sqrt(g * 1e-5))
    assert li.stable_timestep(st, topo, 0.85, np.float64) == pytest.approx(expect, rel=1e-14)


def test_coupled_ledger_balance_closes():
    """The parity_runner.py:337-345 balance identity over 60 coupled steps with the
    local-inertial surface (all cells active, every drainage path)."""
    inp = small_network_scenario(seed=17, h_init=0.2)
    inp = inp._replace(surface=inp.surface._replace(active=np.ones_like(inp.surface.active)))
    io = to_oracle(inp, np.float64)
    s, c, p, r, st, f = io
    area = s.cell_area
    init = np.sum(st.h * area) + np.sum(st.node_volume) + np.sum(st.recharge_storage)
    led = orc.zero_ledger()
    t = 0.0
    for _ in range(60):
        dt = 0.5
        st, d = oracle_step(io, st, dt, rain=orc.rain_rate(f, t), q=orc.external_q(f, t + 0.25),
                            dtype=np.float64, surface_mode="local_inertial")
        for k, v in d.items():
            led[k] += v
        t += dt
    final = np.sum(st.h * area) + np.sum(st.node_volume) + np.sum(st.recharge_storage)
    lhs = init + led["rainfall_input"] + led["external_boundary_inflow"]
    rhs = (final + led["infiltration_loss"] + led["outfall_discharge"] + led["unreturned_overflow"]
           + led["recharge_loss"] - led["boundary_net_flux"])
    assert abs(lhs - rhs) / lhs < 1e-12
    assert led["boundary_outflow"] > 0 and led["surface_capture"] > 0
