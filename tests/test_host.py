"""CPU-side checks: the C ABI library loads and exports every declared symbol,
host topology preparation equals the reference's edge builder, synthetic
workloads have the configured sizes, and the backend fails closed without a GPU."""

import re
from pathlib import Path

import numpy as np
import pytest

from floodhip import _lib, synth
from floodhip.topology import build_surface_topology
from oracle import topology_ref

ROOT = Path(__file__).resolve().parents[1]


def _ensure_built():
    if not _lib.LIB_PATH.exists():
        from floodhip import build
        build.build()


def test_library_exports_every_header_symbol():
    _ensure_built()
    header = (ROOT / "include" / "floodhip.h").read_text()
    declared = set(re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations parsed"
    import ctypes
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [name for name in sorted(declared) if not hasattr(lib, name)]
    assert not missing, missing
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)


def test_default_params_are_the_reference_constants():
    _ensure_built()
    import ctypes
    lib = _lib.load()
    p = _lib.FhParams()
    lib.fh_default_params(ctypes.byref(p))
    assert p.gravity == 9.80665 and p.min_depth == 1e-5 and p.cfl == 0.85
    assert p.cfl_interval == 32 and p.reuse_factor == 0.8 and p.max_dt == 1.0 and p.dt_floor == 1e-3
    assert p.max_speed == 5.0 and p.inlet_orifice_q == 0.62 * 0.04 and p.slot_wave_speed == 30.0


def test_param_structs_match_the_header_layout(tmp_path):
    """The ctypes mirrors of fh_params / fh_clock have the C layout: every field
    at the offset gcc gives it, and the same total size."""
    import ctypes
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = []
    for struct, cls in (("fh_params", _lib.FhParams), ("fh_clock", _lib.FhClock)):
        lines.append(f'printf("{struct} size %zu\\n", sizeof({struct}));')
        for name, _ in cls._fields_:
            lines.append(f'printf("{struct} {name} %zu\\n", offsetof({struct}, {name}));')
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "floodhip.h"\nint main(void) {\n'
                   + "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(w[:2]): int(w[2]) for w in (ln.split() for ln in out) if len(w) == 3}
    for struct, cls in (("fh_params", _lib.FhParams), ("fh_clock", _lib.FhClock)):
        assert got[(struct, "size")] == ctypes.sizeof(cls), struct
        for name, _ in cls._fields_:
            assert got[(struct, name)] == getattr(cls, name).offset, (struct, name)
    _ensure_built()
    p = _lib.FhParams()
    _lib.load().fh_default_params(ctypes.byref(p))
    assert p.share_mesh == 1 and p.surface_mode == 0 and p.cell_order == 1


def test_fails_closed_without_gpu():
    _ensure_built()
    lib = _lib.load()
    if lib.fh_device_count() > 0:
        pytest.skip("a GPU is visible")
    from floodhip.errors import BackendUnavailableError
    from floodhip.solver import HipParitySolver
    inp = synth.make_plain_small(n=4, extent=40.0)
    with pytest.raises(BackendUnavailableError):
        HipParitySolver(inp.surface)
    assert lib.fh_create(None) is None
    assert b"no HIP device" in lib.fh_global_error()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_vectorized_topology_equals_reference_edge_builder(seed):
    rng = np.random.default_rng(seed)
    nodes, tri = synth.jittered_delaunay(9, 7, 90.0, 70.0, rng)
    tri = tri.copy()
    flip = rng.random(tri.shape[0]) < 0.5        # random orientations exercise the CCW fix
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    bed = rng.random(tri.shape[0]).astype(np.float32)
    active = rng.random(tri.shape[0]) > 0.2
    got, _ = build_surface_topology(nodes, tri, bed, 0.03, active_mask=active, boundary_mode="outflow")
    ref = topology_ref.build_surface_loop(nodes, tri, bed, 0.03, active=active, boundary_code_value=4)
    for k in ("edge_left", "edge_right", "edge_nx", "edge_ny", "edge_length", "boundary_code", "cell_area",
              "characteristic_length"):
        assert np.array_equal(getattr(got, k), ref[k]), k


def test_unit_square_topology_by_hand():
    nodes = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 2.0], [0.0, 2.0]])
    tri = np.array([[0, 2, 1], [0, 2, 3]])      # first is clockwise
    s, t = build_surface_topology(nodes, tri, np.zeros(2, np.float32), 0.03, boundary_mode="closed")
    assert np.allclose(s.cell_area, 2.0)
    assert s.edge_left.shape[0] == 5 and (s.edge_right >= 0).sum() == 1
    e = int(np.flatnonzero(s.edge_right >= 0)[0])
    assert np.isclose(s.edge_length[e], np.sqrt(8.0))
    # normal of the diagonal points out of its left cell
    L = s.edge_left[e]
    cen = nodes[t[L]].mean(0)
    mid = np.array([1.0, 1.0])
    assert np.dot(np.array([s.edge_nx[e], s.edge_ny[e]]), mid - cen) > 0
    assert np.allclose(s.characteristic_length, 2 * 2.0 / (4.0 + np.sqrt(8.0)))
    assert set(s.boundary_code[s.edge_right < 0]) == {0}


def test_workload_sizes():
    a = synth.make_plain_small()
    assert a.surface.cell_area.shape[0] == 20_000 and a.surface.edge_left.shape[0] == 30_200
    assert (a.surface.boundary_code[a.surface.edge_right < 0] == 4).all()
    assert np.isclose(float(a.forcing.rainfall_rate_mps[0]), 50.0 / 3.6e6, rtol=1e-6)
    b = synth.make_catchment(20_000, seed=5, rain_hours=1)
    n = b.surface.cell_area.shape[0]
    assert 19_000 < n < 21_000
    nn = b.coupling.node_storage_area.shape[0]
    assert nn == round(0.12 * n) and b.pipe.link_from.shape[0] == round(1.177 * nn)
    assert b.coupling.inlet_cell.max() < n and b.pipe.link_to.max() < nn


def test_delaunay_count_matches_euler():
    rng = np.random.default_rng(3)
    pts, tri = synth.jittered_delaunay(30, 20, 300.0, 200.0, rng, extra_points=17)
    assert tri.shape[0] == 2 * pts.shape[0] - 2 * (30 + 20) - 2
