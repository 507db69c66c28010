"""The functional entry points (advance_parity_step, stable_timestep: the
reference's v1_parity.py:430-469 and :153-158) on the GPU against the oracle,
looped as a caller of the reference's functional API loops them: the context
stays resident across calls, and a topology changed in place is picked up.

Tolerances as in the coupled step tests (float32): per step |dh| <= 1e-5 m,
node volume within 1e-4 of max(V, 1 m^3), ledger terms within 1e-4 relative;
the CFL time is bit-identical (same float32 operations, exact minimum)."""

import numpy as np
import pytest

from floodhip import solver as fs
from helpers import oracle_step, small_network_scenario, to_oracle
from oracle import v1_parity_np as orc

pytestmark = pytest.mark.gpu


def test_looped_advance_parity_step_matches_oracle():
    inp = small_network_scenario(seed=11)
    io = to_oracle(inp)
    f = inp.forcing
    st_ref = io[4]
    st = inp.initial_state
    t = 0.0
    fs.clear_context_cache()
    seen = set()
    for i in range(12):
        dt = 0.4 if i % 2 else 0.25
        rate = orc.rain_rate(io[5], t)
        q = orc.external_q(io[5], t + 0.5 * dt)
        st_ref, rdiag = oracle_step(io, st_ref, dt, rain=rate, q=q)
        st, diag = fs.advance_parity_step(
            st, inp.surface, inp.coupling, inp.pipe, inp.recharge, rainfall_rate_mps=rate,
            rainfall_multiplier=f.rainfall_multiplier, horton_f0_mps=f.horton_f0_mps,
            horton_fc_mps=f.horton_fc_mps, horton_k_per_s=f.horton_k_per_s,
            external_receiver_cell=f.external_receiver_cell, external_q_m3ps=q, dt=dt)
        seen.add(id(fs._contexts._entries[next(reversed(fs._contexts._entries))][2]))
        t += dt
        assert np.max(np.abs(st.h - st_ref.h)) <= 1e-5, i
        vs = np.maximum(np.abs(st_ref.node_volume), 1.0)
        assert np.max(np.abs(st.node_volume - st_ref.node_volume) / vs) <= 1e-4, i
        for k in orc.LEDGER_FIELDS:
            assert abs(getattr(diag, k) - rdiag[k]) <= 1e-4 * max(abs(rdiag[k]), 1e-3), (i, k)
    assert len(seen) == 1  # one resident context served every call
    fs.clear_context_cache()


def test_stable_timestep_follows_the_state_and_in_place_changes():
    inp = small_network_scenario(seed=3)
    io = to_oracle(inp)
    fs.clear_context_cache()
    st = inp.initial_state
    for scale in (1.0, 2.0, 0.5):
        s2 = st._replace(h=(st.h * np.float32(scale)).astype(np.float32),
                         hu=(st.h * np.float32(0.2 * scale)).astype(np.float32))
        o2 = io[4]._replace(h=s2.h, hu=s2.hu)
        assert fs.stable_timestep(s2, inp.surface) == orc.stable_timestep(o2, io[0])
    # change the cell length scale in place: the cached context must not be reused
    inp.surface.characteristic_length[:] *= np.float32(0.5)
    io2 = to_oracle(inp)
    assert fs.stable_timestep(st, inp.surface) == orc.stable_timestep(io[4], io2[0])
    fs.clear_context_cache()
