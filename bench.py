"""Benchmark: triangle-steps/s and %HBM roofline of the coupled solver step on one MI355X.

Workload (N=1): the full production-scale plain method -- 981,880-triangle seeded
Delaunay mesh with 1,474,220 edges, fractal bed, 5 land-cover classes, Horton
infiltration, and the 118,768-node / 139,798-link / 61,317-inlet / 46-outfall
drainage network; float32 state, FP64 ledger and clock, CFL 0.85 policy.
A step is one solver timestep (surface flux + cell update + sources + drainage).
Inputs are resident in HBM before the timed region; the timed region is exactly
`--steps` steps of fh_advance_to (the call blocks until the GPU is done).

Prints one JSON line (stdout).  Progress goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for _p in (str(ROOT), str(ROOT / "gpu-urban-flood-cupy-jax_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
REFERENCE_CUPY_STEPS_PER_S = 758.443          # published, RTX PRO 4000 Blackwell, same mesh size


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_flux_update_981880.md"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel factor);
    None when the summary is absent."""
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 5 and cells[0] == kernel:
                return float(cells[4])
    except (OSError, ValueError):
        return None
    return None


def algorithmic_bytes(kernel, E, C, compensated=False):
    """Algorithmic bytes per launch (float32): each required element once per step."""
    if kernel == "flux":
        return 44.0 * E + 20.0 * C
    if kernel == "update":
        return 20.0 * E + 72.0 * C + (24.0 * C if compensated else 0.0)
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(inputs, steps, numpy_steps):
    """The CPU oracle on a bounded sample of the same workload, from its initial
    state under the CFL policy: the C restatement on every host thread OpenMP
    gets (OMP_NUM_THREADS, else all cores), and the NumPy oracle in one process
    beside it.  Returns (C-oracle rate, threads, seconds, NumPy rate, seconds)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from helpers import to_oracle
    from oracle import v1_parity_c as oc
    from oracle import v1_parity_np as orc
    io = to_oracle(inputs)
    C = inputs.surface.cell_area.shape[0]
    threads = oc.set_threads(0)
    co = oc.COracle(*io[:4], io[5])
    t0 = time.perf_counter()
    _, clock, _ = co.advance_to_time(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 3600.0, 3600.0,
                                     max_steps=steps)
    el = time.perf_counter() - t0
    rate = C * clock.global_step / el
    np_rate, np_el = None, 0.0
    if numpy_steps > 0:
        t0 = time.perf_counter()
        _, c2, _ = orc.advance_to_time(io[4], orc.OracleClock(0.0, 1.0, 0), orc.zero_ledger(), 3600.0, 3600.0,
                                       *io[:4], io[5], max_steps=numpy_steps)
        np_el = time.perf_counter() - t0
        np_rate = C * c2.global_step / np_el
    return rate, threads, el, np_rate, np_el


def _ensemble_rate(members, steps, warmup, opts):
    from floodhip.solver import HipEnsemble
    sizes = [int(m.surface.cell_area.shape[0]) for m in members]
    with HipEnsemble(members, **opts) as ens:
        ens.advance_to(3600.0, 3600.0, max_steps=max(warmup, 1), raise_on_budget=False)
        before = [c["global_step"] for c, _ in ens.member_clocks()]
        t0 = time.perf_counter()
        ens.advance_to(3600.0, 3600.0, max_steps=steps, raise_on_budget=False)
        el = time.perf_counter() - t0
        after = [c["global_step"] for c, _ in ens.member_clocks()]
        prof = ens.profile_kernels(steps=5)
    tri = sum(n * (a - b) for n, a, b in zip(sizes, after, before))
    return tri / el, el, sum(a - b for a, b in zip(after, before)), prof


def batch_leg(base, args, plain_rate):
    """16 members on the production mesh, one launch per stage per step for all of them."""
    from floodhip import synth
    from floodhip.autotune import autotune
    members = synth.make_shared_mesh_ensemble(base, n_members=args.batch_members)
    opts = dict(compensated=True, use_graph=not args.no_graph)
    tuned = None
    if args.batch_options:
        bs, ept, order = (int(v) for v in args.batch_options.split(","))
        opts.update(block_size=bs, edges_per_thread=ept, pack_order=order)
    elif not args.no_autotune:
        # tuned on the whole batch: a shared-mesh batch's block order and cache
        # reuse depend on its member count, so a sample of members mispredicts
        best, table = autotune(members, steps=32, warmup=8, log=log, compensated=True)
        opts.update(best)
        tuned = {"best": best, "table": [[o, r] for o, r in table], "sample_members": len(members)}
    log(f"batched ensemble: {len(members)} members, options {opts}")
    rate, el, steps, prof = _ensemble_rate(members, args.batch_steps, 20, opts)
    # sequential plain baseline with the same state precision: one member alone, default launch options
    seq_rate, _, _, _ = _ensemble_rate([members[0]], args.batch_steps, 20,
                                       dict(compensated=True, use_graph=not args.no_graph))
    log(f"batched: {rate:.4e} tri-steps/s over {steps} member-steps in {el:.3f}s; sequential plain "
        f"{seq_rate:.4e} ({rate / seq_rate:.2f}x)")
    # accuracy of the same batch against its FP64 run: 120 fixed 0.5 s steps
    from floodhip.precision import ensemble_vs_fp64
    fp64 = ensemble_vs_fp64(members, steps=120, dt=0.5, use_graph=True)
    log(f"batched vs FP64 (120 x 0.5 s): worst stored-volume rel {fp64['worst']['stored_volume_rel']:.3e}, "
        f"surface-volume rel {fp64['worst']['surface_volume_rel']:.3e}, depth Linf {fp64['worst']['depth_linf_m']:.3e} m")
    C = int(base.surface.cell_area.shape[0])
    E = int(base.surface.edge_left.shape[0])
    M = len(members)
    # algorithmic bytes of one launch over the batch; the members share the mesh, so
    # the edge tables (20 B per edge, flux) and the cell->edge table (12 B per cell,
    # update) count once, not once per member
    shared = {"flux": 20.0 * E, "update": 12.0 * C}
    kern = {}
    for k in prof:
        b = algorithmic_bytes(k["kernel"], E * M, C * M, True)
        if b:
            b -= (M - 1) * shared.get(k["kernel"], 0.0)
        kern[k["kernel"]] = {"avg_us": round(1e3 * k["avg_ms"], 3), "algorithmic_bytes": b,
                             "GBps": round(b / (k["avg_ms"] * 1e-3) / 1e9, 1) if b else None}
    up = kern.get("update", {})
    roof = ({"bound": "hbm", "kernel": "update", "achieved": up["GBps"], "peak": 8000.0, "unit": "GB/s",
             "frac": up["GBps"] / 8000.0} if up.get("GBps") else None)
    return {"workload": f"{M} members x {C} triangles (shared mesh and network, own rain totals 40-200 mm/h "
                        f"and Horton factors), float32 compensated state, FP64 ledgers, hipGraph",
            "members": M, "value": rate, "unit": "triangle-steps/s", "timed_member_steps": steps,
            "seconds": el, "sequential_plain_value": seq_rate, "speedup_vs_sequential_plain": rate / seq_rate,
            "options": opts, "autotune": tuned, "kernels": kern, "roofline": roof,
            "vs_fp64": {"fixed_dt_s": fp64["fixed_dt_s"], "steps": fp64["steps"], "worst_member": fp64["worst"]}}


def wet_regime_leg(inp, args):
    """The same workload from a uniform 10 cm start, where every cell is wet and the
    CFL dt is about 0.05 s (the regime of the one-hour gate event): ms per step and
    triangle-steps/s over a window of steps inside the event."""
    from floodhip.solver import HipParitySolver
    C = int(inp.surface.cell_area.shape[0])
    st = inp.initial_state._replace(h=np.full(C, 0.1, np.float32))
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, st,
                         use_graph=not args.no_graph) as s:
        s.advance_to(3600.0, 3600.0, max_steps=max(args.warmup, 1), raise_on_budget=False)
        c0 = s.clock()
        t0 = time.perf_counter()
        c1, _ = s.advance_to(3600.0, 3600.0, max_steps=args.wet_steps, raise_on_budget=False)
        el = time.perf_counter() - t0
    steps = c1["global_step"] - c0["global_step"]
    sim = c1["time_s"] - c0["time_s"]
    log(f"wet regime: {steps} steps in {el:.3f}s: {1e3 * el / steps:.4f} ms/step, mean dt {sim / steps:.4f}s")
    return {"workload": "the same 981,880-triangle coupled workload from a 10 cm wet start (every cell wet)",
            "value": C * steps / el, "unit": "triangle-steps/s", "steps": steps, "ms_per_step": 1e3 * el / steps,
            "mean_dt_s": sim / steps, "simulated_s_per_wall_s": sim / el}


def local_inertial_leg(inp, args):
    """The reference's production surface mode (surface_step_mode: local_inertial,
    engine.py:219-256) on the same workload and network: the step's CFL pass, the
    edge discharges, the donor limiting and the update, then the same drainage."""
    from floodhip.solver import HipParitySolver
    C = int(inp.surface.cell_area.shape[0])
    s = HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                        surface_mode="local_inertial", use_graph=not args.no_graph)
    final_t = 3600.0
    s.advance_to(final_t, final_t, max_steps=max(args.warmup, 1), raise_on_budget=False)
    c0 = s.clock()
    t0 = time.perf_counter()
    c1, _ = s.advance_to(final_t, final_t, max_steps=args.li_steps, raise_on_budget=False)
    el = time.perf_counter() - t0
    steps = c1["global_step"] - c0["global_step"]
    sim = c1["time_s"] - c0["time_s"]
    st = s.get_state()
    ok = bool(np.isfinite(st.h).all() and st.h.min() >= -1e-7)
    prof = s.profile_kernels(steps=args.profile_steps)
    s.close()
    log(f"local-inertial: {steps} steps in {el:.3f}s: {1e3 * el / steps:.4f} ms/step, "
        f"{C * steps / el:.4e} tri-steps/s, {sim / el:.1f} simulated s per s (mean dt {sim / steps:.3f}s)")
    for k in prof:
        log(f"kernel {k['kernel']:>18s}: {1e3 * k['avg_ms']:8.2f} us  x{k['launches']}")
    return {"workload": "the same 981,880-triangle coupled workload, surface_mode local_inertial, float32, hipGraph",
            "value": C * steps / el, "unit": "triangle-steps/s", "steps": steps, "ms_per_step": 1e3 * el / steps,
            "simulated_s_per_wall_s": sim / el, "mean_dt_s": sim / steps, "state_check": ok,
            "kernels": {k["kernel"]: {"avg_us": round(1e3 * k["avg_ms"], 3)} for k in prof}}


def batch_uneven_leg(args):
    from floodhip import synth
    from floodhip.autotune import autotune
    log("generating the 48-member uneven ensemble")
    members = synth.make_uneven_ensemble()
    sizes = [int(m.surface.cell_area.shape[0]) for m in members]
    best, table = autotune(members, steps=16, warmup=4, log=log, compensated=True)
    opts = dict(compensated=True, **best)
    rate, el, steps, _ = _ensemble_rate(members, 64, 8, opts)
    seq = []
    for m in members:
        r, _, _, _ = _ensemble_rate([m], 64, 8, dict(compensated=True))
        seq.append((int(m.surface.cell_area.shape[0]), r))
    # sequential plain: total work / total time if each member ran alone at its own rate
    seq_rate = sum(n for n, _ in seq) / sum(n / r for n, r in seq)
    log(f"uneven batched {rate:.4e} vs sequential plain {seq_rate:.4e} (all {len(members)} members)")
    return {"members": len(members), "triangles": int(sum(sizes)), "value": rate, "unit": "triangle-steps/s",
            "options": opts, "sequential_plain_value": seq_rate, "sequential_plain_members": len(members),
            "speedup_vs_sequential_plain": rate / seq_rate}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=1000, help="steps of the C oracle (all host threads)")
    ap.add_argument("--numpy-steps", type=int, default=60, help="steps of the NumPy oracle (one process)")
    ap.add_argument("--wet-steps", type=int, default=2000)
    ap.add_argument("--no-wet", action="store_true", help="skip the wet-regime leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--compensated", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--block-size", type=int, default=64)
    ap.add_argument("--no-remap", action="store_true")
    ap.add_argument("--batch-members", type=int, default=16)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batch-options", default=None,
                    help="block size, edges per thread, pack order of the shared-mesh batch (e.g. 256,2,0) "
                         "instead of autotuning them")
    ap.add_argument("--no-uneven", action="store_true", help="skip the 48-member uneven ensemble")
    ap.add_argument("--no-li", action="store_true", help="skip the local-inertial surface mode leg")
    ap.add_argument("--li-steps", type=int, default=1000)
    ap.add_argument("--copy-probe-mib", type=int, default=256,
                    help="size of the streaming-copy probe (achievable HBM bandwidth); 0 skips it")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: every run here is single-GPU "
              "(the multi-GPU sharding is a design, not a measured path). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from floodhip import synth
    from floodhip.solver import HipParitySolver

    log("generating the 981,880-triangle coupled workload")
    t0 = time.perf_counter()
    inp = synth.make_production_scale()
    log(f"generated in {time.perf_counter() - t0:.1f}s")
    C = int(inp.surface.cell_area.shape[0])
    E = int(inp.surface.edge_left.shape[0])
    s = HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                        compensated=args.compensated, use_graph=not args.no_graph,
                        block_size=args.block_size, xcd_remap=not args.no_remap)
    final_t = 3600.0
    s.advance_to(final_t, final_t, max_steps=max(args.warmup, 1), raise_on_budget=False)
    log(f"resident on the GPU: {s.device_bytes() / 2**20:.1f} MiB")
    log(f"warmup {args.warmup} steps done, t={s.clock()['time_s']:.2f}s")
    c0 = s.clock()
    t0 = time.perf_counter()
    c1, ledger = s.advance_to(final_t, final_t, max_steps=args.steps, raise_on_budget=False)
    el = time.perf_counter() - t0
    steps = c1["global_step"] - c0["global_step"]
    if steps != args.steps:
        log(f"warning: only {steps} steps before t={final_t}")
    tri_steps = C * steps / el
    log(f"timed {steps} steps in {el:.3f}s: {1e3 * el / steps:.4f} ms/step, {tri_steps:.4e} tri-steps/s, "
        f"t={c1['time_s']:.2f}s")
    st = s.get_state()
    if not (np.isfinite(st.h).all() and st.h.min() >= -1e-7):
        log("state check FAILED")
        sys.exit(3)
    prof = s.profile_kernels(steps=args.profile_steps)
    for k in prof:
        log(f"kernel {k['kernel']:>18s}: {1e3 * k['avg_ms']:8.2f} us  x{k['launches']}")
    by_time = sorted([k for k in prof if k["kernel"] in ("flux", "update")], key=lambda k: -k["avg_ms"])
    dom = by_time[0]
    ab = algorithmic_bytes(dom["kernel"], E, C, args.compensated)
    achieved = ab / (dom["avg_ms"] * 1e-3) / 1e9
    kernels = {k["kernel"]: {"avg_us": round(1e3 * k["avg_ms"], 3),
                             "algorithmic_bytes": algorithmic_bytes(k["kernel"], E, C, args.compensated),
                             "GBps": (round(algorithmic_bytes(k["kernel"], E, C, args.compensated)
                                            / (k["avg_ms"] * 1e-3) / 1e9, 1)
                                      if algorithmic_bytes(k["kernel"], E, C, args.compensated) else None)}
               for k in prof}
    s.close()
    traffic = None if args.compensated else measured_traffic(dom["kernel"])
    roofline = {"bound": "hbm", "kernel": dom["kernel"], "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "algorithmic_bytes": ab, "avg_us": 1e3 * dom["avg_ms"],
                "byte_model": "each required array element once per step (flux 44E + 20C, update 20E + 72C)"}
    if traffic:
        # the same kernel's HBM bytes from the rocprofv3 counters (committed summary of
        # an earlier run of this workload), over this run's kernel time
        roofline.update({"traffic_source": str(TRAFFIC_FILE.relative_to(ROOT)) + " (rocprofv3 FETCH_SIZE and "
                         "WRITE_SIZE per launch, scaled by the copy kernel's counts)",
                         "frac_counter_bytes": traffic / (dom["avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS})
    # the plain run's accuracy against its FP64 run (120 fixed 0.5 s steps)
    from floodhip.precision import ensemble_vs_fp64
    fp64 = ensemble_vs_fp64([inp], steps=120, dt=0.5, compensated=args.compensated, use_graph=True)
    log(f"plain vs FP64 (120 x 0.5 s): stored-volume rel {fp64['worst']['stored_volume_rel']:.3e}, "
        f"depth Linf {fp64['worst']['depth_linf_m']:.3e} m")
    wet = None if args.no_wet else wet_regime_leg(inp, args)
    local_inertial = None if args.no_li else local_inertial_leg(inp, args)
    batched = None
    if not args.no_batch and args.batch_members > 0:
        batched = batch_leg(inp, args, tri_steps)
    uneven = None if (args.no_uneven or args.no_batch) else batch_uneven_leg(args)
    copy = None
    if args.copy_probe_mib > 0:
        # achievable bandwidth: a streaming copy, one 16-byte element per thread, with
        # ordinary and with non-temporal loads and stores
        import ctypes
        from floodhip import _lib
        nf = args.copy_probe_mib * 2**20 // 4
        copy = {"bytes_per_launch": 2 * nf * 4}
        for key, nt in (("ordinary", 0), ("nontemporal", 1)):
            ms = ctypes.c_double()
            if _lib.load().fh_copy_probe(nf, 20, nt, ctypes.byref(ms)) == 0:
                copy[key] = {"avg_us": 1e3 * ms.value, "GBps": 2 * nf * 4 / (ms.value * 1e-3) / 1e9}
                log(f"copy probe ({key}): {copy[key]['GBps']:.1f} GB/s "
                    f"({copy['bytes_per_launch'] / 2**20:.0f} MiB per launch)")
        best = max((v["GBps"] for k, v in copy.items() if isinstance(v, dict)), default=None)
        if best:
            copy["GBps"] = best
            roofline["frac_of_best_copy"] = achieved / best
    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        log(f"CPU baseline: oracle, {args.cpu_steps} steps of the same workload")
        v, threads, cel, npv, npel = cpu_baseline(inp, args.cpu_steps, args.numpy_steps)
        cpu = {"value": v, "unit": "triangle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the 981,880-triangle coupled workload from its initial state "
                         f"under the CFL policy (C restatement of the oracle, float32, OpenMP, {cel:.1f}s)",
               "host_cpu": _cpu_model(), "host_logical_cpus": os.cpu_count(),
               "numpy_single_process": (None if npv is None else
                                        {"value": npv, "unit": "triangle-steps/s", "cores": 1,
                                         "sample": f"{args.numpy_steps} steps, NumPy float32 oracle ({npel:.1f}s)"})}
        log(f"CPU baseline {v:.4e} tri-steps/s on {threads} threads; NumPy {npv}")
    out = {
        "metric": "triangle-steps/s + %HBM peak, 1 GPU; batched ensemble vs sequential plain",
        "value": tri_steps,
        "unit": "triangle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain, full production scale: 981,880-triangle Delaunay mesh + 118,768-node "
                               "drainage network, float32 state, FP64 ledger/clock",
                   "triangles": C, "edges": E, "nodes": int(inp.coupling.node_storage_area.shape[0]),
                   "links": int(inp.pipe.link_from.shape[0]), "inlets": int(inp.coupling.inlet_cell.shape[0]),
                   "seed": int(inp.report.get("seed", 0)), "compensated": bool(args.compensated),
                   "hip_graph": not args.no_graph, "parallelism": "single GPU"},
        "roofline": roofline,
        "kernels": kernels,
        "simulated_s_per_wall_s": (c1["time_s"] - c0["time_s"]) / el,
        "vs_fp64": {"fixed_dt_s": fp64["fixed_dt_s"], "steps": fp64["steps"], **fp64["members"][0]},
        "wet_regime": wet,
        "local_inertial": local_inertial,
        "copy_probe": copy,
        "batched_shared_mesh": batched,
        "batched_uneven": uneven,
        "cpu_baseline": cpu,
        "reference_published_tri_steps_per_s": REFERENCE_CUPY_STEPS_PER_S * C,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
