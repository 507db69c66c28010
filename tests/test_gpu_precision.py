"""FP64 instantiation of the kernels against the FP64 oracle, and the compensated
FP32 update against both (dam break in a closed frictionless channel).

Tolerances: FP64 GPU vs FP64 oracle max |dh| <= 1e-10 m (both round every
operation in double, FMA contraction off); compensated FP32 must be at least as
close to FP64 as plain FP32 in L2 and conserve volume to 1e-6 relative."""

import numpy as np
import pytest

from floodhip import synth
from floodhip.solver import HipParitySolver
from floodhip.precision import stoker_depth
from helpers import oracle_step, to_oracle
from oracle import v1_parity_np as orc

pytestmark = pytest.mark.gpu


def _dam(nx=400, ny=6):
    return synth.make_dam_break(nx=nx, ny=ny, length=2000.0, width=30.0, h_up=4.0, h_down=0.01)


def _run(inp, precision, compensated, t_end, fixed_dt=None):
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                         precision=precision, compensated=compensated, fixed_dt=fixed_dt) as s:
        clock, _ = s.advance_to(t_end, t_end, max_steps=100_000)
        return s.get_state_f64()[0], clock


def test_fp64_kernels_match_fp64_oracle():
    inp = _dam()
    dt, n = 0.05, 200
    io = to_oracle(inp, np.float64)
    st = io[4]
    for _ in range(n):
        st, _ = oracle_step(io, st, dt, dtype=np.float64)
    h, clock = _run(inp, 64, False, n * dt, fixed_dt=dt)
    assert clock["global_step"] == n
    assert np.max(np.abs(h - st.h)) <= 1e-10


def test_compensated_fp32_is_closer_to_fp64():
    inp = _dam(nx=1000, ny=8)
    t_end = 40.0
    area = inp.surface.cell_area.astype(np.float64)
    v0 = np.sum(inp.initial_state.h.astype(np.float64) * area)
    h64, _ = _run(inp, 64, False, t_end)
    h32, _ = _run(inp, 32, False, t_end)
    h32c, _ = _run(inp, 32, True, t_end)
    l2 = lambda a: float(np.sqrt(np.sum((a - h64) ** 2 * area) / area.sum()))  # noqa: E731
    assert l2(h32c) <= l2(h32) * 1.05
    assert abs(np.sum(h32c * area) - v0) / v0 < 1e-6
    assert abs(np.sum(h64 * area) - v0) / v0 < 1e-12
    # all three resolve the Stoker solution
    exact = stoker_depth(inp.host["centroids"][:, 0], t_end, 4.0, 0.01, 1000.0)
    for h in (h64, h32, h32c):
        assert np.sum(np.abs(h - exact) * area) / area.sum() < 0.1   # numerical diffusion at the front
