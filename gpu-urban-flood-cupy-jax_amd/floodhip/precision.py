"""Precision study: the 200,000-triangle dam break run three ways on the GPU.

* FP32 plain: float32 state, FP64 ledger and clock;
* FP32 compensated: float32 state with a carried TwoSum compensation of the
  per-cell update (flux divergence and source terms), FP64 ledger and clock;
* FP64: the same kernels instantiated for double, the oracle the other two are
  measured against.

Each result reports depth L2 / Linf error against the FP64 run and against
Stoker's analytic solution, and the relative volume drift (closed, frictionless
basin: volume is exactly conserved in exact arithmetic).

Usage: python -m floodhip.precision [--out FILE]
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import synth
from .solver import HipParitySolver


def stoker_depth(x, t, h_l, h_r, x0, g=9.80665):
    """Stoker's wet-bed dam-break depth (1-D), shock state by bisection."""
    c_l = np.sqrt(g * h_l)

    def f(hm):
        um_raref = 2.0 * (c_l - np.sqrt(g * hm))
        s = np.sqrt(0.5 * g * hm * (hm + h_r) / h_r)
        return um_raref - s * (1.0 - h_r / hm)

    lo, hi = h_r * (1 + 1e-12), h_l
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    hm = 0.5 * (lo + hi)
    um = 2.0 * (c_l - np.sqrt(g * hm))
    s = np.sqrt(0.5 * g * hm * (hm + h_r) / h_r)
    xi = (np.asarray(x) - x0) / max(t, 1e-12)
    cm = np.sqrt(g * hm)
    h = np.where(xi <= -c_l, h_l, 0.0)
    h = np.where((xi > -c_l) & (xi <= um - cm), (2 * c_l - xi) ** 2 / (9 * g), h)
    h = np.where((xi > um - cm) & (xi <= s), hm, h)
    return np.where(xi > s, h_r, h)


def run_form(inp, duration, precision, compensated):
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                         precision=precision, compensated=compensated) as s:
        t0 = time.perf_counter()
        clock, _ = s.advance_to(duration, duration, max_steps=1_000_000)
        el = time.perf_counter() - t0
        h, hu, hv = s.get_state_f64()
    return h, clock, el


def study(duration=120.0, inp=None) -> dict:
    inp = inp or synth.make_dam_break()
    area = inp.surface.cell_area.astype(np.float64)
    v0 = float(np.sum(inp.initial_state.h.astype(np.float64) * area))
    x = inp.host["centroids"][:, 0]
    exact = stoker_depth(x, duration, 10.0, 0.001, 5000.0)
    forms = {}
    h64, c64, e64 = run_form(inp, duration, 64, False)
    for name, prec, comp in (("fp64", 64, False), ("fp32_plain", 32, False), ("fp32_compensated", 32, True)):
        if name == "fp64":
            h, clock, el = h64, c64, e64
        else:
            h, clock, el = run_form(inp, duration, prec, comp)
        d = h - h64
        forms[name] = {
            "steps": int(clock["global_step"]), "seconds": el,
            "depth_l2_vs_fp64_m": float(np.sqrt(np.sum(d * d * area) / area.sum())),
            "depth_linf_vs_fp64_m": float(np.max(np.abs(d))),
            "depth_l1_vs_stoker_m": float(np.sum(np.abs(h - exact) * area) / area.sum()),
            "depth_linf_vs_stoker_m": float(np.max(np.abs(h - exact))),
            "relative_volume_drift": float((np.sum(h * area) - v0) / v0),
        }
    return {"workload": "dam break, 10 km x 200 m channel, 200,000 triangles, 10 m | 0.001 m at x = 5 km, "
                        "flat frictionless bed, reflective walls", "duration_s": duration, "forms": forms}


def ensemble_vs_fp64(members, steps: int = 120, dt: float = 0.5, compensated: bool = True, **opts) -> dict:
    """A batch in float32 (with the compensated update unless compensated=False)
    against the same batch in FP64 (the same kernels for double), both with the
    same fixed dt, so every
    member sits at the same simulated time.  Per member: the relative error of
    the stored volume (surface sum h * A plus the node volumes) and of the
    surface volume alone, and the depth L2 / Linf errors.  Returns the worst
    member of each and the per-member lists."""
    from .solver import HipEnsemble
    runs = {}
    for prec in (32, 64):
        with HipEnsemble(members, precision=prec, compensated=compensated and prec == 32, fixed_dt=dt,
                         **opts) as ens:
            ens.advance_to(1.0e9, 1.0e9, max_steps=steps, raise_on_budget=False)
            out = []
            for m in range(ens.members):
                ens.select(m)
                st = ens.get_state()
                h = ens.get_state_f64()[0] if prec == 64 else st.h.astype(np.float64)
                out.append((h, np.asarray(st.node_volume, np.float64), ens.clock()["global_step"]))
            runs[prec] = out
    per = []
    for inp, (h32, v32, n32), (h64, v64, n64) in zip(members, runs[32], runs[64]):
        if n32 != n64:
            raise RuntimeError(f"step counts differ: {n32} vs {n64}")
        area = np.asarray(inp.surface.cell_area, np.float64)
        s32, s64 = float(np.sum(h32 * area)), float(np.sum(h64 * area))
        t32, t64 = s32 + float(np.sum(v32)), s64 + float(np.sum(v64))
        d = h32 - h64
        per.append({"steps": int(n32), "stored_volume_rel": (t32 - t64) / max(abs(t64), 1.0),
                    "surface_volume_rel": (s32 - s64) / max(abs(s64), 1.0),
                    "depth_l2_m": float(np.sqrt(np.mean(d * d))), "depth_linf_m": float(np.max(np.abs(d)))})
    worst = {k: max(abs(p[k]) for p in per)
             for k in ("stored_volume_rel", "surface_volume_rel", "depth_l2_m", "depth_linf_m")}
    return {"fixed_dt_s": dt, "steps": steps, "worst": worst, "members": per}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--duration", type=float, default=120.0)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    res = study(args.duration)
    txt = json.dumps(res, indent=2)
    print(txt)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt + "\n")


if __name__ == "__main__":
    main()
