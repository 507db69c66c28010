"""The HIP path against the golden fixtures (tests/golden/*.npz): stored inputs,
stored step schedule (dt, rain rate, external inflow per step), stored expected
state and summed ledger.  No oracle runs here.

Tolerances: float32 |dh| <= 1e-5 m after 25 steps, node volume 1e-4 relative
(of max(V, 1 m^3)), link flow 1e-4 relative (of max(|Q|, 1e-3 m^3/s)), ledger
1e-4 relative; float64 kernels |dh|, |dhu| <= 1e-10 after 100 steps (node
volumes and link flows cross the boundary as float32: 1e-5 relative), ledger
1e-6 relative.  The same bounds hold for the local-inertial fixtures
(coupled_li_*, 25 float32 steps / 60 float64 steps)."""

import numpy as np
import pytest

from floodhip.solver import HipParitySolver
from helpers import load_golden
from oracle import v1_parity_np as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["coupled_f32", "coupled_f64", "coupled_li_f32", "coupled_li_f64"])
def test_hip_matches_golden(name):
    inp, sched, exp, ledger, dtype = load_golden(name)
    f64 = dtype == np.float64
    got_ledger = dict.fromkeys(orc.LEDGER_FIELDS, 0.0)
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                         precision=64 if f64 else 32, surface_mode=sched["surface_mode"]) as s:
        for dt, rain, q in zip(sched["dt"], sched["rain"], sched["q"]):
            d = s.step(float(dt), float(rain), q if q.size else None)
            for k in orc.LEDGER_FIELDS:
                got_ledger[k] += getattr(d, k)
        got = s.get_state()
        h, hu, _ = s.get_state_f64()
    tol_h = 1e-10 if f64 else 1e-5
    tol_r = 1e-5 if f64 else 1e-4
    assert np.max(np.abs(h - exp.h)) <= tol_h
    assert np.max(np.abs(hu - exp.hu)) <= (tol_h if f64 else 1e-4)
    vs = np.maximum(np.abs(exp.node_volume), 1.0)
    assert np.max(np.abs(got.node_volume - exp.node_volume) / vs) <= tol_r
    qs = np.maximum(np.abs(exp.link_flow), 1e-3)
    assert np.max(np.abs(got.link_flow - exp.link_flow) / qs) <= tol_r
    lt = 1e-6 if f64 else 1e-4
    for k in orc.LEDGER_FIELDS:
        assert abs(got_ledger[k] - ledger[k]) <= lt * max(abs(ledger[k]), 1e-3), (k, got_ledger[k], ledger[k])
