"""Writes the golden fixtures of this folder: seeded coupled scenarios (inputs) and
the oracle's outputs after a fixed sequence of steps (expected values).

The reference holds no arrays for this path (its Gurugram assets and .npz results
are not part of the repository), so these fixtures come from the oracle, whose
functions are pinned by closed forms and invariants in tests/test_oracle_*.py.
They freeze the oracle's answers so that a later change to the oracle, to the
scenario generators or to the device path shows up as a fixture mismatch, and
they let the device tests run against stored numbers.

Each .npz holds every field of the input NamedTuples (``surface.*``,
``coupling.*``, ``pipe.*``, ``recharge.*``, ``state.*``, ``forcing.*``), the step
schedule (``dt``, ``rain``, ``q``: one row per step) and the expected outputs
(``out.<state field>`` and ``ledger.<field>``, the sum of the per-step ledger
terms in float64), and ``surface_mode`` when the case runs the local-inertial
surface step.

    python tests/golden/make_golden.py [NAME ...]   # rewrite the fixtures (all, or the named ones)
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path[:0] = [str(HERE.parent), str(HERE.parents[1]), str(HERE.parents[1] / "gpu-urban-flood-cupy-jax_amd")]

from helpers import oracle_step, small_network_scenario, to_oracle  # noqa: E402
from oracle import v1_parity_np as orc  # noqa: E402

GROUPS = ("surface", "coupling", "pipe", "recharge", "initial_state", "forcing")


def flatten_inputs(inp) -> dict:
    out = {}
    for g in GROUPS:
        tup = getattr(inp, g)
        for k, v in tup._asdict().items():
            out[f"{g}.{k}"] = np.asarray(v)
    return out


def schedule(inp, steps, dt_of, external=True):
    """Per-step (dt, rain rate, external inflow) as the oracle evaluates them."""
    f = to_oracle(inp)[5]
    nrx = np.asarray(inp.forcing.external_receiver_cell).shape[0]
    dts, rains, qs = [], [], []
    t = 0.0
    for i in range(steps):
        dt = dt_of(i)
        dts.append(dt)
        rains.append(float(orc.rain_rate(f, t)))
        qs.append(np.asarray(orc.external_q(f, t + 0.5 * dt), np.float32) if external and nrx
                  else np.zeros(nrx, np.float32))
        t += dt
    return np.asarray(dts, np.float64), np.asarray(rains, np.float64), np.stack(qs).reshape(steps, nrx)


def run_oracle(inp, dts, rains, qs, dtype, surface_mode="full_swe"):
    io = to_oracle(inp, dtype)
    st = io[4]
    ledger = orc.zero_ledger()
    for dt, rain, q in zip(dts, rains, qs):
        st, diag = oracle_step(io, st, float(dt), rain=float(rain), q=np.asarray(q, dtype), dtype=dtype,
                               surface_mode=surface_mode)
        for k in orc.LEDGER_FIELDS:
            ledger[k] += float(diag[k])
    return st, ledger


CASES = {
    # float32 (the reference's precision): every drainage path, external inflow,
    # surcharge entries, all boundary kinds, inactive cells, alternating dt
    "coupled_f32": dict(make=lambda: small_network_scenario(seed=7), steps=25,
                        dt_of=lambda i: 0.3 if i % 2 else 0.45, dtype=np.float32, external=True),
    # float64 kernels against the FP64 oracle, fixed dt
    "coupled_f64": dict(make=lambda: small_network_scenario(seed=21, h_init=0.2, with_external=False),
                        steps=100, dt_of=lambda i: 0.4, dtype=np.float64, external=False),
    # the local-inertial surface step (triangular_backend.py:1181-1291) with the
    # same sources and drainage, float32 and float64
    "coupled_li_f32": dict(make=lambda: small_network_scenario(seed=13, h_init=0.2), steps=25,
                           dt_of=lambda i: 0.3 if i % 2 else 0.45, dtype=np.float32, external=True,
                           surface_mode="local_inertial"),
    "coupled_li_f64": dict(make=lambda: small_network_scenario(seed=29, h_init=0.2, with_external=False),
                           steps=60, dt_of=lambda i: 0.4, dtype=np.float64, external=False,
                           surface_mode="local_inertial"),
}


def build(name):
    case = CASES[name]
    inp = case["make"]()
    dts, rains, qs = schedule(inp, case["steps"], case["dt_of"], case["external"])
    mode = case.get("surface_mode", "full_swe")
    st, ledger = run_oracle(inp, dts, rains, qs, case["dtype"], mode)
    arrays = flatten_inputs(inp)
    arrays.update({"dt": dts, "rain": rains, "q": qs, "dtype": np.asarray(np.dtype(case["dtype"]).name)})
    if mode != "full_swe":
        arrays["surface_mode"] = np.asarray(mode)
    arrays.update({f"out.{k}": np.asarray(v) for k, v in st._asdict().items()})
    arrays.update({f"ledger.{k}": np.float64(v) for k, v in ledger.items()})
    return arrays


def main(names=None):
    for name in names or CASES:
        arrays = build(name)
        path = HERE / f"{name}.npz"
        np.savez_compressed(path, **arrays)
        print(f"{path.name}: {path.stat().st_size} bytes, {len(arrays)} arrays")


if __name__ == "__main__":
    main(sys.argv[1:])
