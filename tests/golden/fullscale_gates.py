"""Full-scale comparison with the CPU oracle: the 981,880-triangle coupled
production workload (118,768 nodes, 139,798 links, 61,317 inlets), float32.

Two procedures, each a CPU half (the C oracle, oracle/v1_parity_c.py, on all
host cores) and a GPU half (the HIP solver), joined by a comparison on the CPU:

1. The one-hour event through the reference's eight acceptance gates
   (hybrid_flood/operational/parity_validation.py:95-192).  The event starts from
   a uniform 10 cm layer: the gates' flood-extent CSI is intersection /
   max(union, 1), which is 0 on a snapshot where neither run has a cell at 5 cm
   or more, so from a dry start the first snapshot (after the first step) fails
   that gate for any pair of runs.  From that start every cell is wet and the CFL
   dt falls to about 0.05 s: the hour takes about 72,700 steps, so both sides run
   with a budget of 400,000 steps per output interval (the reference runner's
   default is 20,000).  The oracle half takes about an hour on 8 cores; it
   checkpoints every 2,000 steps and resumes from the last checkpoint.  With
   --surface-mode local_inertial both halves run the local-inertial step (the C
   restatement covers both modes).

       python tests/golden/fullscale_gates.py oracle-event DIR [--threads N] [--surface-mode M]
       python tests/golden/fullscale_gates.py gpu-event OUT.npz [--surface-mode M]   (GPU)
       python tests/golden/fullscale_gates.py gates DIR/oracle_event.npz OUT.npz SUMMARY.json [--surface-mode M]

2. 300 s of the dry-start event under the CFL policy, in float32 and in FP64 on
   both sides, for the drainage node volumes: FP64 against FP64 separates
   kernel mistakes (which FP64 would not hide) from float32 conditioning.

       python tests/golden/fullscale_gates.py oracle-300 DIR
       python tests/golden/fullscale_gates.py gpu-300 OUT.npz              (GPU)
       python tests/golden/fullscale_gates.py compare-300 DIR OUT.npz SUMMARY.json
       python tests/golden/fullscale_gates.py resync-300 SUMMARY.json      (GPU + C oracle)
       python tests/golden/fullscale_gates.py trace-300 OUT.npz            (GPU + C oracle)

   resync-300 restarts the GPU from the oracle's float32 state before every step
   and compares one step at a time (the one-step error); trace-300 follows both
   FP64 trajectories step by step (depths, node volumes, outfall nodes).  With
   --surface-mode local_inertial the 300 s procedures run the local-inertial
   step (the C restatement covers both modes).
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path[:0] = [str(HERE.parent), str(HERE.parent.parent), str(HERE.parent.parent / "gpu-urban-flood-cupy-jax_amd")]

import numpy as np  # noqa: E402

from floodhip import synth  # noqa: E402
from floodhip.runner import output_targets  # noqa: E402
from helpers import to_oracle  # noqa: E402
from oracle import v1_parity_np as orc  # noqa: E402

DURATION = 3600.0
BUDGET = 400_000          # steps per output interval
CHUNK = 2_000             # oracle steps between checkpoints
EVENT_FIELDS = ("snapshot_times_s", "depth_snapshots_m", "drain_node_volume_snapshots_m3", "steps_at_snapshot")


def event_inputs():
    inp = synth.make_production_scale()
    h0 = np.full(inp.surface.cell_area.shape, 0.1, np.float32)
    return inp._replace(initial_state=inp.initial_state._replace(h=h0))


def _report(inp, state, ledger, depth, steps):
    """The report fields the gates read (parity_runner.py:333-345), FP64 ledger."""
    area = np.asarray(inp.surface.cell_area, np.float64)
    init_s = float(np.sum(np.asarray(inp.initial_state.h, np.float64) * area))
    init_n = float(np.sum(np.asarray(inp.initial_state.node_volume, np.float64)))
    fin_s = float(np.sum(np.asarray(depth[-1], np.float64) * area))
    fin_n = float(np.sum(np.asarray(state.node_volume, np.float64)))
    lhs = init_s + init_n + ledger["rainfall_input"] + ledger["external_boundary_inflow"]
    rhs = (fin_s + fin_n + ledger["infiltration_loss"] + ledger["outfall_discharge"]
           + ledger["unreturned_overflow"] + ledger["recharge_loss"] - ledger["boundary_net_flux"])
    return {"max_depth_m": float(np.max(depth)), **{f"{k}_m3": float(v) for k, v in ledger.items()},
            "mass_residual_ratio": (lhs - rhs) / max(abs(lhs), 1.0), "step_count": int(steps)}


def oracle_event(out_dir: Path, threads: int, duration=DURATION, mode="full_swe"):
    from oracle import v1_parity_c as oc
    out_dir.mkdir(parents=True, exist_ok=True)
    if threads:
        oc.set_threads(threads)
    inp = event_inputs()
    s, c, p, r, st, f = to_oracle(inp)
    co = oc.COracle(s, c, p, r, f, surface_mode=mode)
    targets = [float(np.float32(1.0e-6))] + [float(x) for x in output_targets(duration)[1:]]
    ckpt = out_dir / "checkpoint.npz"
    clock, ledger, snaps, k, steps_in = orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), [], 0, 0
    if ckpt.exists():
        with np.load(ckpt, allow_pickle=False) as z:
            st = orc.ParityState(*[z[f"state_{n}"] for n in orc.ParityState._fields])
            clock = orc.OracleClock(float(z["time_s"]), float(z["cached_dt_s"]), int(z["global_step"]))
            ledger = json.loads(str(z["ledger"]))
            k, steps_in = int(z["target_index"]), int(z["steps_in_interval"])
            snaps = json.loads(str(z["snapshots"]))
            depth_done = [z[f"depth_{i}"] for i in range(len(snaps))]
            nodes_done = [z[f"nodes_{i}"] for i in range(len(snaps))]
        print(f"resumed at step {clock.global_step}, t = {clock.time_s:.3f} s", flush=True)
    else:
        depth_done, nodes_done = [], []
    t0 = time.time()
    while k < len(targets):
        target = targets[k]
        n = min(CHUNK, BUDGET - steps_in)
        st, c1, ledger = co.advance_to_time(st, clock, ledger, target, duration, max_steps=n)
        steps_in += c1.global_step - clock.global_step
        clock = c1
        if clock.time_s >= target or steps_in >= BUDGET:
            if clock.time_s < target:
                raise RuntimeError(f"step budget reached at t = {clock.time_s:.3f} s before {target} s")
            snaps.append({"time_s": clock.time_s, "step": clock.global_step})
            depth_done.append(np.array(st.h, np.float32))
            nodes_done.append(np.array(st.node_volume, np.float32))
            k, steps_in = k + 1, 0
        arrays = {f"state_{n}": np.asarray(getattr(st, n)) for n in orc.ParityState._fields}
        arrays.update({f"depth_{i}": d for i, d in enumerate(depth_done)})
        arrays.update({f"nodes_{i}": v for i, v in enumerate(nodes_done)})
        tmp = out_dir / "checkpoint.tmp.npz"
        np.savez(tmp, time_s=clock.time_s, cached_dt_s=clock.cached_dt_s, global_step=clock.global_step,
                 ledger=json.dumps(ledger), target_index=k, steps_in_interval=steps_in,
                 snapshots=json.dumps(snaps), **arrays)
        tmp.replace(ckpt)
        rate = (time.time() - t0) / max(1, clock.global_step)
        print(f"step {clock.global_step}: t = {clock.time_s:.3f} s, dt cached {clock.cached_dt_s:.4f} s "
              f"({rate * 1e3:.1f} ms/step this session)", flush=True)
    depth = np.stack(depth_done)
    record = {"snapshot_times_s": np.asarray([x["time_s"] for x in snaps], np.float32),
              "steps_at_snapshot": np.asarray([x["step"] for x in snaps], np.int64),
              "depth_snapshots_m": depth, "drain_node_volume_snapshots_m3": np.stack(nodes_done)}
    report = _report(inp, st, ledger, depth, clock.global_step)
    np.savez_compressed(out_dir / "oracle_event.npz", report=json.dumps(report), **record)
    print(json.dumps(report), flush=True)


def gpu_event(out: Path, duration=DURATION, mode="full_swe"):
    from floodhip.runner import run_hip_v1_parity
    inp = event_inputs()
    t0 = time.time()
    res = run_hip_v1_parity(inp, duration_s=duration, max_steps_per_hour=BUDGET, keep_final_state=False,
                            surface_mode=mode)
    el = time.time() - t0
    rep = res["report"]
    rep["wall_s"] = el
    np.savez_compressed(out, report=json.dumps(rep, default=float), **{k: res[k] for k in EVENT_FIELDS})
    print(json.dumps({"steps": rep["step_count"], "wall_s": el, "max_depth_m": rep["max_depth_m"]}), flush=True)


def _load_record(path):
    with np.load(path, allow_pickle=False) as z:
        rec = {k: z[k] for k in z.files if k != "report"}
        rec["report"] = json.loads(str(z["report"]))
    return rec


def gates(oracle_npz: Path, gpu_npz: Path, out_json: Path, mode="full_swe"):
    from floodhip import gates as g
    ref, cand = _load_record(oracle_npz), _load_record(gpu_npz)
    inp = event_inputs()
    res = g.validate(ref, cand, inp.surface.cell_area)
    d = cand["depth_snapshots_m"][-1].astype(np.float64) - ref["depth_snapshots_m"][-1]
    rel_v = (np.abs(cand["drain_node_volume_snapshots_m3"][-1].astype(np.float64)
                    - ref["drain_node_volume_snapshots_m3"][-1]) /
             np.maximum(np.abs(ref["drain_node_volume_snapshots_m3"][-1]), 1.0))
    summary = {"workload": "981,880-triangle coupled production workload, 10 cm wet start, one hour, float32; "
                           "reference: the CPU oracle (C restatement, all host cores)", "surface_mode": mode,
               "passed": bool(res["passed"]), "gates": res["gates"], "summary": res["summary"],
               "steps": {"gpu": int(cand["report"]["step_count"]), "oracle": int(ref["report"]["step_count"])},
               "last_snapshot": {"max_abs_dh_m": float(np.max(np.abs(d))),
                                 "cells_dh_gt_1mm": int(np.sum(np.abs(d) > 1e-3)),
                                 "node_volume_rel_max": float(rel_v.max()),
                                 "nodes_rel_gt_1e-3": int(np.sum(rel_v > 1e-3))},
               "ledger_comparison": res["ledger_comparison"], "ledger_terms_missing": res["ledger_terms_missing"],
               "hourly": res["hourly"]}
    out_json.write_text(json.dumps(summary, indent=1, default=float) + "\n")
    print(json.dumps({k: summary[k] for k in ("passed", "gates", "steps", "last_snapshot")}, default=float))


def _dry_inputs():
    return synth.make_production_scale()


def _tag(dtype_tag: str, mode: str) -> str:
    return dtype_tag if mode == "full_swe" else f"li_{dtype_tag}"


class _Oracle:
    """The oracle's CFL-policy advance: the C restatement on the host cores, in
    either surface mode (tests/test_oracle_c.py pins it against the NumPy oracle
    step by step in both)."""

    def __init__(self, io, dtype, mode):
        from oracle import v1_parity_c as oc
        self.io, self.dtype, self.mode = io, dtype, mode
        s, c, p, r, _, f = io
        self.co = oc.COracle(s, c, p, r, f, dtype=dtype, surface_mode=mode)

    def advance(self, st, clock, ledger, t_end, max_steps):
        return self.co.advance_to_time(st, clock, ledger, t_end, DURATION, max_steps=max_steps)


def oracle_300(out_dir: Path, threads: int, t_end=300.0, mode="full_swe"):
    from oracle import v1_parity_c as oc
    out_dir.mkdir(parents=True, exist_ok=True)
    if threads:
        oc.set_threads(threads)
    inp = _dry_inputs()
    for dtype, tag in ((np.float32, "f32"), (np.float64, "f64")):
        io = to_oracle(inp, dtype)
        t0 = time.time()
        st, clock, led = _Oracle(io, dtype, mode).advance(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(),
                                                          t_end, 100_000)
        print(f"oracle {_tag(tag, mode)}: {clock.global_step} steps in {time.time() - t0:.1f} s", flush=True)
        np.savez_compressed(out_dir / f"oracle_300_{_tag(tag, mode)}.npz", h=st.h, node_volume=st.node_volume,
                            link_flow=st.link_flow, steps=clock.global_step, time_s=clock.time_s,
                            ledger=json.dumps(led))


def gpu_300(out: Path, t_end=300.0, mode="full_swe"):
    from floodhip.solver import HipParitySolver
    inp = _dry_inputs()
    arrays = {}
    for prec in (32, 64):
        with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                             precision=prec, surface_mode=mode) as s:
            c, led = s.advance_to(t_end, DURATION, max_steps=100_000)
            st = s.get_state()
            h = s.get_state_f64()[0] if prec == 64 else st.h
        arrays.update({f"h_{prec}": h, f"node_volume_{prec}": st.node_volume, f"link_flow_{prec}": st.link_flow,
                       f"steps_{prec}": c["global_step"], f"ledger_{prec}": json.dumps(led._asdict())})
        print(f"gpu {prec}: {c['global_step']} steps", flush=True)
    np.savez_compressed(out, **arrays)


def compare_300(oracle_dir: Path, gpu_npz: Path, out_json: Path, mode="full_swe"):
    inp = _dry_inputs()
    g = np.load(gpu_npz, allow_pickle=False)
    out = {"workload": "981,880-triangle coupled production workload, dry start, CFL policy to t = 300 s",
           "surface_mode": mode}
    for prec, tag in ((32, "f32"), (64, "f64")):
        o = np.load(oracle_dir / f"oracle_300_{_tag(tag, mode)}.npz", allow_pickle=False)
        vg = g[f"node_volume_{prec}"].astype(np.float64)
        vo = o["node_volume"].astype(np.float64)
        rel = np.abs(vg - vo) / np.maximum(np.abs(vo), 1.0)
        worst = np.argsort(rel)[::-1][:12]
        dh = np.asarray(g[f"h_{prec}"], np.float64) - o["h"].astype(np.float64)
        lg, lo = json.loads(str(g[f"ledger_{prec}"])), json.loads(str(o["ledger"]))
        out[tag] = {
            "steps": [int(g[f"steps_{prec}"]), int(o["steps"])],
            "depth_max_abs_m": float(np.max(np.abs(dh))),
            "depth_rmse_m": float(np.sqrt(np.mean(dh ** 2))),
            "node_volume_rel_max": float(rel.max()), "nodes_rel_gt_1e-3": int(np.sum(rel > 1e-3)),
            "nodes_rel_gt_1e-6": int(np.sum(rel > 1e-6)),
            "node_volume_abs_max_m3": float(np.max(np.abs(vg - vo))),
            "stored_node_volume_rel": float((vg.sum() - vo.sum()) / vo.sum()),
            "worst_nodes": [{"node": int(i), "oracle_m3": float(vo[i]), "gpu_m3": float(vg[i]),
                             "capacity_m3": float(inp.coupling.node_storage_area[i]
                                                  * inp.coupling.node_max_depth[i]),
                             "outfall": bool(inp.pipe.outfall_area[i] > 0)} for i in worst],
            "ledger_rel": {k: (lg[k] - lo[k]) / max(abs(lo[k]), 1e-3) for k in lo},
        }
    out_json.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps({k: {kk: v[kk] for kk in ("steps", "depth_max_abs_m", "node_volume_rel_max",
                                               "nodes_rel_gt_1e-3", "node_volume_abs_max_m3")}
                      for k, v in out.items() if k in ("f32", "f64")}))


def resync_300(out_json: Path, t_end=300.0, mode="full_swe"):
    """Float32 one-step error along the oracle's trajectory: before every step the
    GPU takes the oracle's state (float32, exact), then both take the same step
    (the oracle's dt and rain rate); the step results are compared node by node.
    A kernel that computes the oracle's step differs only by transcendental
    rounding in every step, however far the two free-running trajectories drift."""
    from floodhip.solver import HipParitySolver
    inp = _dry_inputs()
    io = to_oracle(inp)
    s_, c_, p_, r_, st, f_ = io
    outfall = np.asarray(inp.pipe.outfall_area) > 0
    rows = []
    oracle = _Oracle(io, np.float32, mode)
    clock, ledger = orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger()
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                         surface_mode=mode) as s:
        while clock.time_s < t_end:
            before = st
            st, c1, ledger = oracle.advance(before, clock, ledger, t_end, 1)
            if clock.global_step % 20 == 0:
                print(f"resync step {clock.global_step}, t = {clock.time_s:.2f} s", flush=True)
            dt = c1.time_s - clock.time_s
            rate = orc.rain_rate(f_, clock.time_s)
            s.set_state(before)
            s.step(dt, rate)
            got = s.get_state()
            rel = np.abs(got.node_volume.astype(np.float64) - st.node_volume) / np.maximum(np.abs(st.node_volume), 1.0)
            qrel = np.abs(got.link_flow.astype(np.float64) - st.link_flow) / np.maximum(np.abs(st.link_flow), 1e-2)
            rows.append({"step": clock.global_step, "dt": dt, "dh_max": float(np.max(np.abs(got.h - st.h))),
                         "node_rel_max": float(rel.max()), "outfall_node_rel_max": float(rel[outfall].max()),
                         "link_rel_max": float(qrel.max())})
            clock = c1
    worst = {k: max(r[k] for r in rows) for k in ("dh_max", "node_rel_max", "outfall_node_rel_max", "link_rel_max")}
    out = {"workload": "981,880-triangle coupled production workload, dry start, oracle trajectory to t = 300 s; "
                       "the GPU restarts every step from the oracle's float32 state",
           "surface_mode": mode, "steps": len(rows), "worst_one_step_error": worst, "per_step": rows}
    out_json.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps({"steps": len(rows), **worst}), flush=True)


def trace_300(out_npz: Path, t_end=300.0, mode="full_swe"):
    """FP64 on both sides, free-running under the CFL policy: after every step the
    largest depth difference, the largest relative node-volume difference and the
    node volumes of every outfall node, with both dts, to show where and how fast
    the trajectories part."""
    from floodhip.solver import HipParitySolver
    inp = _dry_inputs()
    io = to_oracle(inp, np.float64)
    st = io[4]
    oracle = _Oracle(io, np.float64, mode)
    outfall = np.flatnonzero(np.asarray(inp.pipe.outfall_area) > 0)
    clock, ledger = orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger()
    vo, vg, relmax, dhmax, dts = [], [], [], [], []
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                         precision=64, surface_mode=mode) as s:
        while clock.time_s < t_end:
            st, c1, ledger = oracle.advance(st, clock, ledger, t_end, 1)
            gc, _ = s.advance_to(t_end, DURATION, max_steps=1, raise_on_budget=False)
            v, _ = s.get_drainage_f64()
            h = s.get_state_f64()[0]
            vo.append(np.asarray(st.node_volume)[outfall])
            vg.append(v[outfall])
            relmax.append(float(np.max(np.abs(v - st.node_volume) / np.maximum(np.abs(st.node_volume), 1.0))))
            dhmax.append(float(np.max(np.abs(h - st.h))))
            dts.append((c1.time_s - clock.time_s, gc["last_dt_s"]))
            if clock.global_step % 20 == 0:
                print(f"trace step {clock.global_step}: max |dh| {dhmax[-1]:.3e} m, node rel {relmax[-1]:.3e}",
                      flush=True)
            clock = c1
    np.savez_compressed(out_npz, outfall_nodes=outfall, oracle=np.stack(vo), gpu=np.stack(vg),
                        node_rel_max=np.asarray(relmax), depth_max_abs=np.asarray(dhmax), dt=np.asarray(dts))
    thr = (1e-13, 1e-10, 1e-7, 1e-4, 1e-2)
    first = {t: next((i for i, x in enumerate(relmax) if x > t), None) for t in thr}
    first_h = {t: next((i for i, x in enumerate(dhmax) if x > t), None) for t in thr}
    dt_diff = next((i for i, (a, b) in enumerate(dts) if a != b), None)
    print(json.dumps({"steps": len(relmax), "node_first_step_above": first, "depth_first_step_above": first_h,
                      "final_node_rel_max": relmax[-1], "final_depth_max_abs": dhmax[-1],
                      "dt_first_differs_at_step": dt_diff}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("mode", choices=["oracle-event", "gpu-event", "gates", "oracle-300", "gpu-300", "compare-300",
                                     "resync-300", "trace-300"])
    ap.add_argument("paths", nargs="+")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--duration", type=float, default=DURATION)
    ap.add_argument("--surface-mode", default="full_swe", choices=["full_swe", "local_inertial"],
                    help="surface operator of the runs (the gates mode compares what it is given)")
    a = ap.parse_args()
    P = [Path(x) for x in a.paths]
    if a.mode == "oracle-event":
        oracle_event(P[0], a.threads, a.duration, mode=a.surface_mode)
    elif a.mode == "gpu-event":
        gpu_event(P[0], a.duration, mode=a.surface_mode)
    elif a.mode == "gates":
        gates(P[0], P[1], P[2], mode=a.surface_mode)
    elif a.mode == "oracle-300":
        oracle_300(P[0], a.threads, mode=a.surface_mode)
    elif a.mode == "gpu-300":
        gpu_300(P[0], mode=a.surface_mode)
    elif a.mode == "compare-300":
        compare_300(P[0], P[1], P[2], mode=a.surface_mode)
    elif a.mode == "resync-300":
        resync_300(P[0], mode=a.surface_mode)
    else:
        trace_300(P[0], mode=a.surface_mode)


if __name__ == "__main__":
    main()
