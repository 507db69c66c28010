"""Pins the CPU oracle's surface step (v1_parity.py:110-256) against closed forms,
conservation laws, symmetry and a textbook HLL restated in a rotated frame."""

import numpy as np
import pytest

from helpers import basin, oracle_step, stoker_depth, to_oracle
from oracle import v1_parity_np as orc

G = orc.GRAVITY


def run(inp, steps, dt, dtype=np.float64, rain=0.0):
    io = to_oracle(inp, dtype)
    st = io[4]
    led = orc.zero_ledger()
    for _ in range(steps):
        st, d = oracle_step(io, st, dt, rain=rain, dtype=dtype)
        for k, v in d.items():
            led[k] += v
    return st, led, io


def test_lake_at_rest_over_bumpy_bed_is_well_balanced():
    # closed walls; eta = h + z exactly constant in FP64.  (With code-4 outflow
    # walls the reference zeroes the whole flux vector, pressure included, when
    # rounding noise makes the mass flux point inward -- v1_parity.py:191-193 --
    # so rest is only preserved there when the noise is exactly zero.)
    bump = lambda c: 1.5 * np.exp(-((c - 60.0) ** 2).sum(1) / 400.0) + 0.002 * c[:, 0]
    inp = basin(n=10, code="closed", bed=bump)
    io = list(to_oracle(inp, np.float64))
    h0 = 3.0 - io[0].bed
    io[4] = io[4]._replace(h=h0)
    st = io[4]
    for _ in range(50):
        st, _ = oracle_step(io, st, 0.2, dtype=np.float64)
    # residual motion comes only from float32 edge normals/lengths (sum n L != 0 at
    # ~1e-7); without the hydrostatic pressure correction the bump drives ~0.1 m/s
    assert np.max(np.abs(st.h - h0)) < 1e-6
    assert np.max(np.abs(st.hu)) < 1e-5 and np.max(np.abs(st.hv)) < 1e-5


def test_closed_basin_conserves_volume_f64_and_f32():
    hump = lambda c: 0.5 + 0.4 * np.exp(-((c - 40.0) ** 2).sum(1) / 300.0)
    inp = basin(n=12, code="closed", h0=hump, manning=0.02)
    area = inp.surface.cell_area.astype(np.float64)
    v0 = np.sum(inp.initial_state.h.astype(np.float64) * area)
    st64, _, _ = run(inp, 80, 0.1, np.float64)
    assert abs(np.sum(st64.h * area) - v0) / v0 < 1e-12
    st32, _, _ = run(inp, 80, 0.1, np.float32)
    assert abs(np.sum(st32.h.astype(np.float64) * area) - v0) / v0 < 2e-6
    assert np.max(np.abs(st32.h - st64.h)) < 1e-4   # FP32 vs FP64 oracle


def test_mirror_symmetry_about_diagonal():
    # split-square diagonals run along x = y, so the mesh maps to itself under x <-> y
    hump = lambda c: 0.3 + 0.6 * np.exp(-((c[:, 0] - 30.0) ** 2 + (c[:, 1] - 30.0) ** 2) / 200.0)
    inp = basin(n=10, code="reflective", h0=hump, manning=0.01)
    st, _, _ = run(inp, 40, 0.1, np.float64)
    cen = inp.host["centroids"]
    key = {(round(x, 6), round(y, 6)): i for i, (x, y) in enumerate(cen)}
    mirror = np.array([key[(round(y, 6), round(x, 6))] for x, y in cen])
    assert np.max(np.abs(st.h - st.h[mirror])) < 1e-10
    assert np.max(np.abs(st.hu - st.hv[mirror])) < 1e-10


def test_rain_only_closed_flat_basin_gives_r_times_t():
    inp = basin(n=6, code="closed")
    r = 50.0 / 3.6e6
    st, led, _ = run(inp, 100, 0.5, np.float64, rain=r)
    assert np.max(np.abs(st.h - r * 50.0)) < 1e-12 * r * 50.0 * 10
    area = inp.surface.cell_area.astype(np.float64).sum()
    assert abs(led["rainfall_input"] - r * 50.0 * area) / (r * 50.0 * area) < 1e-12


def test_horton_infiltration_matches_geometric_series():
    inp = basin(n=4, code="closed", h0=lambda c: np.full(c.shape[0], 0.2))
    C = inp.surface.cell_area.shape[0]
    f0, fc, k = 2e-5, 5e-6, 1.0 / 600.0
    inp = inp._replace(forcing=inp.forcing._replace(horton_f0_mps=np.full(C, f0, np.float32),
                                                    horton_fc_mps=np.full(C, fc, np.float32),
                                                    horton_k_per_s=np.full(C, k, np.float32)))
    dt, n = 1.0, 300
    st, led, io = run(inp, n, dt, np.float64)
    f0d, fcd, kd = (float(np.float32(v)) for v in (f0, fc, k))
    q = np.exp(-kd * dt)
    expected = fcd * n * dt + (f0d - fcd) * dt * (1 - q ** n) / (1 - q)
    h0 = float(np.float32(0.2))
    area = inp.surface.cell_area.astype(np.float64)
    # volume: exact geometric series (closed basin, every cell wet every step)
    assert abs(led["infiltration_loss"] - expected * area.sum()) < 1e-12 * area.sum()
    assert abs(np.sum((h0 - st.h) * area) - expected * area.sum()) < 1e-12 * area.sum()
    # per cell: up to the tiny float32-geometry flows of a flat basin
    assert np.max(np.abs((h0 - st.h) - expected)) < 1e-8
    assert np.allclose(st.wet_time_s, n * dt)


def test_dry_domain_has_zero_flux_and_no_motion():
    inp = basin(n=5, code="outflow", bed=lambda c: 0.01 * c[:, 0])
    st, led, _ = run(inp, 5, 1.0, np.float32)
    assert np.all(st.h == 0) and np.all(st.hu == 0) and led["boundary_outflow"] == 0.0


def _hll_rotated(hl, ul, vl, hr, ur, vr, n, g=G, eps=1e-5):
    """Textbook HLL in the edge frame (normal/tangential velocity), rotated back."""
    nx, ny = n
    unl, utl = ul * nx + vl * ny, -ul * ny + vl * nx
    unr, utr = ur * nx + vr * ny, -ur * ny + vr * nx
    cl, cr = np.sqrt(g * hl), np.sqrt(g * hr)
    sl, sr = min(unl - cl, unr - cr), max(unl + cl, unr + cr)
    Fl = np.array([hl * unl, hl * unl ** 2 + 0.5 * g * hl ** 2, hl * unl * utl])
    Fr = np.array([hr * unr, hr * unr ** 2 + 0.5 * g * hr ** 2, hr * unr * utr])
    Ql, Qr = np.array([hl, hl * unl, hl * utl]), np.array([hr, hr * unr, hr * utr])
    if sl >= 0:
        F = Fl
    elif sr <= 0:
        F = Fr
    else:
        F = (sr * Fl - sl * Fr + sl * sr * (Qr - Ql)) / (sr - sl)
    return np.array([F[0], F[1] * nx - F[2] * ny, F[1] * ny + F[2] * nx])


@pytest.mark.parametrize("state", [(1.0, 0.3, -0.1, 0.5, -0.2, 0.4), (2.0, 1.5, 0.0, 1.9, 1.4, 0.2),
                                   (0.2, -3.0, 0.5, 0.1, -2.5, 0.0)])
def test_flat_bed_edge_flux_matches_textbook_hll(state):
    hl, ul, vl, hr, ur, vr = state
    nx, ny = 0.6, 0.8
    z = np.zeros(1)
    fl, fr = orc._hydrostatic_flux_pair(np.array([hl]), np.array([hl * ul]), np.array([hl * vl]), z,
                                        np.array([hr]), np.array([hr * ur]), np.array([hr * vr]), z,
                                        np.array([nx]), np.array([ny]), np.float64)
    ref = _hll_rotated(hl, ul, vl, hr, ur, vr, (nx, ny))
    assert np.allclose(fl[:, 0], ref, rtol=1e-12, atol=1e-12)
    assert np.allclose(fr[:, 0], ref, rtol=1e-12, atol=1e-12)   # flat bed: no pressure correction


def test_reflective_wall_has_zero_mass_flux():
    inp = basin(n=6, code="closed", h0=lambda c: 0.5 + 0.2 * (c[:, 0] > 30))
    st, led, _ = run(inp, 20, 0.2, np.float64)
    assert abs(led["boundary_inflow"]) < 1e-12 and abs(led["boundary_outflow"]) < 1e-12


def test_two_cell_step_by_hand():
    # two triangles forming the unit square; left wet 1.0 at rest, right 0.5, flat bed, closed walls
    from floodhip.topology import build_surface_topology
    nodes = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    tri = np.array([[0, 1, 2], [0, 2, 3]])
    surface, _ = build_surface_topology(nodes, tri, np.zeros(2, np.float32), 0.0, boundary_mode="closed")
    s = orc.cast_tuple(orc.SurfaceTopology(*surface), np.float64)
    st = orc.ParityState(np.array([1.0, 0.5]), np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(0),
                         np.zeros(0), np.zeros(0))
    dt = 0.01
    new, bnet, bin_, bout = orc.surface_flux_step(st, s, dt, np.float64)
    # shared diagonal from (0,0) to (1,1): length sqrt(2), normal from cell 0 to cell 1 is (-1, 1)/sqrt(2)
    e = int(np.flatnonzero(s.edge_right >= 0)[0])
    n = np.array([s.edge_nx[e], s.edge_ny[e]])
    L = np.sqrt(2.0)
    Fm = _hll_rotated(1.0, 0, 0, 0.5, 0, 0, n)[0]   # mass flux left -> right
    hl = 1.0 - dt * Fm * L / 0.5
    hr = 0.5 + dt * Fm * L / 0.5
    left = int(s.edge_left[e])
    assert np.isclose(new.h[left], hl if left == 0 else hr, rtol=1e-13)
    assert np.isclose(new.h[1 - left], hr if left == 0 else hl, rtol=1e-13)
    assert abs(bnet) < 1e-15


def test_dam_break_matches_stoker():
    # 1-D channel: 200 x 2 squares, h 1.0 | 0.1, reflective, frictionless, f64
    from floodhip import synth
    from floodhip.topology import build_surface_topology, cell_centroids
    nodes, tri = synth.split_square_grid(1, 1.0)
    nx, ny, Lx = 400, 2, 400.0
    xs = np.linspace(0, Lx, nx + 1)
    ys = np.linspace(0, 2.0, ny + 1)
    X, Y = np.meshgrid(xs, ys)
    nodes = np.column_stack((X.ravel(), Y.ravel()))
    tris = []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            tris += [[a, a + 1, a + nx + 2], [a, a + nx + 2, a + nx + 1]]
    tri = np.array(tris)
    surface, tri = build_surface_topology(nodes, tri, np.zeros(len(tri), np.float32), 0.0,
                                          boundary_mode="closed")
    cen = cell_centroids(nodes, tri)
    s = orc.cast_tuple(orc.SurfaceTopology(*surface), np.float64)
    h = np.where(cen[:, 0] < 200.0, 1.0, 0.1)
    st = orc.ParityState(h, np.zeros_like(h), np.zeros_like(h), np.zeros_like(h), np.zeros(0), np.zeros(0),
                         np.zeros(0))
    t, dt = 0.0, 0.05
    while t < 30.0 - 1e-9:
        st, *_ = orc.surface_flux_step(st, s, dt, np.float64)
        t += dt
    exact = stoker_depth(cen[:, 0], t, 1.0, 0.1, 200.0)
    err = np.mean(np.abs(st.h - exact))
    assert err < 0.01, err
    # the plateau depth between rarefaction and shock is captured
    mid = (cen[:, 0] > 215) & (cen[:, 0] < 235)
    assert abs(np.median(st.h[mid]) - np.median(exact[mid])) < 0.01


def test_velocity_clamp_caps_speed_and_keeps_direction():
    """The clamp fixes |u| = max_speed where the speed exceeds it (direction kept),
    leaves slower cells bit-identical and zeroes the momentum of dry cells."""
    h = np.array([2.0, 1.0, 0.5, 1e-6, 0.3], np.float64)
    u = np.array([12.0, 3.0, -4.0, 7.0, 0.0])
    v = np.array([-9.0, 4.0, 3.0, 1.0, 6.0])
    st = orc.ParityState(h, h * u, h * v, np.zeros(5), np.zeros(0), np.zeros(0), np.zeros(0))
    out = orc.clamp_velocity(st, maximum_speed=5.0, dtype=np.float64)
    speed = np.hypot(out.hu, out.hv) / np.maximum(h, 1e-12)
    assert speed[0] == pytest.approx(5.0, rel=1e-14)                 # 15 m/s capped to 5
    assert np.arctan2(out.hv[0], out.hu[0]) == pytest.approx(np.arctan2(-9.0, 12.0), rel=1e-14)
    assert out.hu[1] == st.hu[1] and out.hv[1] == st.hv[1]            # exactly 5 m/s: untouched
    assert out.hu[2] == st.hu[2] and out.hv[2] == st.hv[2]            # 5 m/s: untouched
    assert out.hu[3] == 0.0 and out.hv[3] == 0.0                      # below the wet threshold
    assert speed[4] == pytest.approx(5.0, rel=1e-14)                  # 6 m/s capped


def test_stable_timestep_closed_forms():
    """Still water of depth h: dt = cfl * min(Lc) / sqrt(g h).  A moving cell:
    dt = cfl * Lc / (|u| + sqrt(g h)); the minimum picks the binding cell."""
    g = orc.GRAVITY
    lc = np.array([4.0, 10.0, 2.5], np.float64)
    topo = types_surface_with_lc(lc)
    h = np.full(3, 2.0)
    st = orc.ParityState(h, np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(0), np.zeros(0), np.zeros(0))
    assert orc.stable_timestep(st, topo, 0.85, np.float64) == pytest.approx(0.85 * 2.5 / np.sqrt(g * 2.0), rel=1e-14)
    # cell 1 moves at 30 m/s (3-4-5 direction): lc 10 / (30 + c) can bind only if smaller
    hu = np.array([0.0, 2.0 * 18.0, 0.0]); hv = np.array([0.0, 2.0 * 24.0, 0.0])
    st = st._replace(hu=hu, hv=hv)
    expect = 0.85 * min(4.0 / np.sqrt(g * 2.0), 10.0 / (30.0 + np.sqrt(g * 2.0)), 2.5 / np.sqrt(g * 2.0))
    assert orc.stable_timestep(st, topo, 0.85, np.float64) == pytest.approx(expect, rel=1e-14)


def types_surface_with_lc(lc):
    """A minimal SurfaceTopology carrying only what stable_timestep reads."""
    n = lc.shape[0]
    fields = {f: np.zeros(0) for f in orc.SurfaceTopology._fields}
    fields["characteristic_length"] = lc
    fields["cell_area"] = np.ones(n)
    return orc.SurfaceTopology(**fields)
