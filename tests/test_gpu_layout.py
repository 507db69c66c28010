"""Device layout: the breadth-first cell/edge order is a storage permutation
only.  Every cell and edge does the same arithmetic in the same order (each
cell still sums its edges in the caller's edge order), so states must equal the
as-given layout bit for bit; the FP64 ledgers sum the same terms in another
block grouping and agree to rounding."""

import numpy as np
import pytest

from floodhip import synth
from floodhip.solver import HipParitySolver
from helpers import small_network_scenario

pytestmark = pytest.mark.gpu


def _run(inp, cell_order, t_end=60.0, **kw):
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                         cell_order=cell_order, **kw) as s:
        clock, ledger = s.advance_to(t_end, 600.0, max_steps=100_000)
        return s.get_state(), clock, np.asarray(ledger)


@pytest.mark.parametrize("make", [
    lambda: small_network_scenario(seed=11, h_init=0.25),
    lambda: synth.make_catchment(8000, seed=synth.BASE_SEED + 31, rain_hours=1, rain_mm_h=[120.0]),
])
@pytest.mark.parametrize("compensated", [False, True])
def test_locality_order_is_a_pure_permutation(make, compensated):
    inp = make()
    a, ca, la = _run(inp, 0, compensated=compensated)
    b, cb, lb = _run(inp, 1, compensated=compensated)
    assert ca["global_step"] == cb["global_step"]
    assert ca["time_s"] == cb["time_s"]
    for name in ("h", "hu", "hv", "wet_time_s", "node_volume", "link_flow", "recharge_storage"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    assert np.allclose(la, lb, rtol=1e-12, atol=1e-12)


def test_state_roundtrip_through_permuted_layout():
    inp = synth.make_catchment(5000, seed=synth.BASE_SEED + 32)
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                         cell_order=1) as s:
        rng = np.random.default_rng(4)
        C = inp.surface.cell_area.shape[0]
        st = s.get_state()
        h = rng.uniform(0.0, 1.0, C).astype(np.float32)
        hu = rng.uniform(-0.1, 0.1, C).astype(np.float32)
        s.set_state(st._replace(h=h, hu=hu, wet_time_s=np.arange(C, dtype=np.float32)))
        back = s.get_state()
    assert np.array_equal(back.h, h)
    assert np.array_equal(back.hu, hu)
    assert np.array_equal(back.wet_time_s, np.arange(C, dtype=np.float32))


@pytest.mark.parametrize("compensated", [False, True])
def test_source_classes_equal_full_arrays(compensated):
    """The 16-bit source classes hold the same (n^2, m, f0, fc, k) values as the full
    per-cell arrays: identical states and ledgers."""
    inp = synth.make_catchment(8000, seed=synth.BASE_SEED + 34, rain_hours=1, rain_mm_h=[100.0])
    a, ca, la = _run(inp, 1, compensated=compensated, source_classes=0)
    b, cb, lb = _run(inp, 1, compensated=compensated, source_classes=1)
    assert ca["global_step"] == cb["global_step"]
    for name in ("h", "hu", "hv", "wet_time_s", "node_volume", "link_flow", "recharge_storage"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    assert np.array_equal(la, lb)
