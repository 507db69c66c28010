"""Pins the oracle's Dirichlet stage ghost at code-3 edges (V1 semantics,
gurugram_flood/kernels/triangular_backend.py:1294-1327 and
triangular_swe_cuda_kernels.py:86-93): depth max(stage - z_left, 0), zero
momentum, the left cell's bed."""

import numpy as np

from helpers import basin, oracle_step, to_oracle
from oracle import v1_parity_np as orc
from test_oracle_surface import _hll_rotated


def _stage_basin(n=8, bed=None, h0=None, codes=None):
    inp = basin(n=n, extent=80.0, code="stage", bed=bed, h0=h0)
    if codes is not None:
        code = inp.surface.boundary_code.copy()
        b = np.flatnonzero(inp.surface.edge_right < 0)
        code[b] = codes(b.size)
        inp = inp._replace(surface=inp.surface._replace(boundary_code=code))
    return inp


def _run(io, steps, dt, stage, dtype=np.float64):
    st = io[4]
    led = orc.zero_ledger()
    for _ in range(steps):
        st, d = oracle_step(io, st, dt, dtype=dtype, boundary_stage_m=stage)
        for k, v in d.items():
            led[k] += v
    return st, led


def test_first_step_boundary_flux_is_textbook_hll_against_the_stage_ghost():
    """Uniform moving water on a flat bed at z0; half of the boundary edges are stage
    edges, the rest reflective walls.  The boundary volume of the first step is the
    textbook HLL mass flux of (h, u, v) against (stage - z0, 0, 0) at every stage
    edge; walls carry none."""
    # the stage equals the water level: the ghost differs from the left state only by
    # its zero momentum, so water leaves where u.n > 0 and enters where u.n < 0, at
    # a rate a copy ghost (u kept) would not give
    z0, h, u, v, stage = 0.5, 1.0, 0.4, -0.25, 1.5
    inp = _stage_basin(bed=lambda c: np.full(c.shape[0], z0), h0=lambda c: np.full(c.shape[0], h),
                       codes=lambda k: np.where(np.arange(k) % 2 == 0, 3, 0))
    io = list(to_oracle(inp, np.float64))
    io[4] = io[4]._replace(hu=io[4].h * u, hv=io[4].h * v)
    s = io[0]
    dt = 0.05
    _, led = _run(io, 1, dt, stage)
    edges = np.flatnonzero((s.edge_right < 0) & (s.boundary_code == 3))
    hs = stage - float(s.bed[0])
    outward = [_hll_rotated(h, u, v, hs, 0.0, 0.0, (float(s.edge_nx[e]), float(s.edge_ny[e])))[0]
               * float(s.edge_length[e]) * dt for e in edges]
    expect_in = sum(max(-o, 0.0) for o in outward)
    expect_out = sum(max(o, 0.0) for o in outward)
    assert expect_in > 0 and expect_out > 0
    assert abs(led["boundary_inflow"] - expect_in) <= 1e-9 * expect_in
    assert abs(led["boundary_outflow"] - expect_out) <= 1e-9 * expect_out


def test_lake_at_stage_level_over_a_bumpy_bed_stays_at_rest():
    bump = lambda c: 0.8 * np.exp(-((c - 40.0) ** 2).sum(1) / 300.0) + 0.004 * c[:, 0]
    inp = _stage_basin(n=10, bed=bump)
    io = list(to_oracle(inp, np.float64))
    stage = 2.0
    h0 = stage - io[0].bed
    io[4] = io[4]._replace(h=h0)
    st, led = _run(io, 50, 0.2, stage)
    assert np.max(np.abs(st.h - h0)) < 1e-6
    assert np.max(np.abs(st.hu)) < 1e-5 and np.max(np.abs(st.hv)) < 1e-5
    assert abs(led["boundary_net_flux"]) < 1e-6 * float(np.sum(h0 * io[0].cell_area))


def test_stage_fills_and_drains_and_the_ledger_closes():
    area = None
    for stage, sign in ((1.5, +1), (0.2, -1)):
        inp = _stage_basin(h0=lambda c: np.full(c.shape[0], 1.0))
        io = to_oracle(inp, np.float64)
        area = io[0].cell_area
        v0 = float(np.sum(io[4].h * area))
        st, led = _run(io, 200, 0.1, stage)
        v1 = float(np.sum(st.h * area))
        assert sign * (v1 - v0) > 0.05 * v0
        # the only source is the boundary: volume change = net boundary volume
        assert abs((v1 - v0) - led["boundary_net_flux"]) <= 1e-10 * v0
        # the level moves toward the stage
        assert sign * (float(np.mean(st.h)) - 1.0) > 0
    # a stage below the bed is a dry ghost: the edge still lets water out
    inp = _stage_basin(bed=lambda c: np.full(c.shape[0], 3.0), h0=lambda c: np.full(c.shape[0], 0.5))
    _, led = _run(to_oracle(inp, np.float64), 20, 0.1, 1.0)
    assert led["boundary_outflow"] > 0 and led["boundary_inflow"] == 0


def test_without_stage_values_code_3_copies_the_left_state():
    """v1_parity.py has no stage values: its ghost at code 3 is the left state, so a
    lake below any stage level stays put."""
    inp = _stage_basin(h0=lambda c: np.full(c.shape[0], 1.0))
    io = to_oracle(inp, np.float64)
    st, led = _run(io, 20, 0.1, None)
    # (float32 edge geometry leaves ~1e-9 m of motion, as in any closed lake)
    v0 = float(np.sum(io[4].h * io[0].cell_area))
    assert np.max(np.abs(st.h - 1.0)) < 1e-7 and abs(led["boundary_net_flux"]) < 1e-8 * v0
    # per-edge stage values: only the edges given a stage above the lake let water in
    E = io[0].edge_right.shape[0]
    stage = np.full(E, 1.0)
    b = np.flatnonzero(io[0].edge_right < 0)
    stage[b[:5]] = 1.4
    st, led = _run(io, 1, 0.1, stage)
    assert led["boundary_inflow"] > 0 and led["boundary_outflow"] < 1e-12
    touched = np.unique(io[0].edge_left[b[:5]])
    others = np.setdiff1d(np.arange(st.h.size), touched)
    assert np.all(st.h[touched] > 1.0) and np.max(np.abs(st.h[others] - 1.0)) < 1e-12


def test_stage_at_an_inactive_cell_flows_in_and_is_removed():
    """The V1 step masks an inactive cell's state to zero but still computes the flux
    of its boundary edges (every edge that is not an active interior edge is a
    boundary edge, v1_parity.py:170-214): a stage above the cell's bed gives the
    textbook HLL inflow of (0, 0, 0) against (stage - z, 0, 0), counted as boundary
    inflow, and the cell's new depth is then zeroed.  Without stage values such an
    edge carries nothing."""
    z0, stage, dt = 0.5, 1.3, 0.05
    inp = _stage_basin(bed=lambda c: np.full(c.shape[0], z0), h0=lambda c: np.full(c.shape[0], 0.2))
    s = inp.surface
    b = np.flatnonzero(s.edge_right < 0)
    dead_cells = np.unique(s.edge_left[b[::7]])
    active = s.active.copy()
    active[dead_cells] = 0
    inp = inp._replace(surface=s._replace(active=active))
    io = list(to_oracle(inp, np.float64))
    s = io[0]
    st, led = _run(io, 1, dt, stage)
    dead_edges = np.flatnonzero((s.edge_right < 0) & (s.active[s.edge_left] < 0.5))
    assert dead_edges.size >= 2
    inflow_dead = sum(-_hll_rotated(0.0, 0.0, 0.0, stage - z0, 0.0, 0.0,
                                    (float(s.edge_nx[e]), float(s.edge_ny[e])))[0]
                      * float(s.edge_length[e]) * dt for e in dead_edges)
    assert inflow_dead > 0
    # the active cells' boundary edges: the same stage against the depth at rest
    live = np.flatnonzero((s.edge_right < 0) & (s.active[s.edge_left] > 0.5))
    h0 = float(io[4].h[0])      # float32 0.2
    inflow_live = sum(-_hll_rotated(h0, 0.0, 0.0, stage - z0, 0.0, 0.0,
                                    (float(s.edge_nx[e]), float(s.edge_ny[e])))[0]
                      * float(s.edge_length[e]) * dt for e in live)
    assert abs(led["boundary_inflow"] - (inflow_dead + inflow_live)) <= 1e-9 * led["boundary_inflow"]
    assert np.all(st.h[dead_cells] == 0)
    _, led0 = _run(io, 1, dt, None)
    assert led0["boundary_inflow"] < 1e-12
