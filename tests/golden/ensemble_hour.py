"""The two batched ensembles over their full simulated hour, as the benchmark
configurations specify them, against each member run alone with the plain
solver (GPU):

* config5: 16 members on the 981,880-triangle production mesh and network, rain
  totals 40-200 mm over the hour, Horton factors perturbed by up to 20 %;
* config4: 48 catchments of 20k-300k triangles (log-uniform), each with its own
  storm (20-200 mm over the hour) and a network of about 12 % of its triangles.

Both run compensated float32 with FP64 ledgers and hipGraph-captured steps, from
a uniform 10 cm layer to t = 3,600 s with the runner's snapshots (after the first
step and at 3,600 s).  The wet start keeps the gates' flood-extent CSI defined:
it is intersection / max(union, 1), 0 on a snapshot where neither run has a
cell at 5 cm.  The batch is one ``floodhip.runner.run_hip_ensemble``; each member
is then run alone by ``run_hip_v1_parity``.  Per member: bit-identical snapshots,
equal step counts, and the reference's eight gates (batch member against its
plain run).  Also the solver seconds of the batch against the sum over the
members run one after another (advance calls only).

    python tests/golden/ensemble_hour.py config5 OUT.json [--options 256,2,0]
    python tests/golden/ensemble_hour.py config4 OUT.json [--options 256,1,2]

--duration S runs S simulated seconds instead of the hour; --batch-only skips the
plain runs and the comparison (for a kernel trace of the batch alone).
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

from floodhip import gates, synth  # noqa: E402
from floodhip.runner import run_hip_ensemble, run_hip_v1_parity  # noqa: E402

DURATION = 3600.0
BUDGET = 400_000      # steps per output interval


def _plain(v):
    return v.item() if isinstance(v, np.generic) else np.asarray(v).tolist()


def members_of(config: str):
    if config == "config5":
        members = synth.make_shared_mesh_ensemble(synth.make_production_scale(), n_members=16)
    else:
        members = synth.make_uneven_ensemble()
    return [m._replace(initial_state=m.initial_state._replace(h=np.full(m.surface.cell_area.shape, 0.1, np.float32)))
            for m in members]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("config", choices=["config5", "config4"])
    ap.add_argument("out", type=Path)
    ap.add_argument("--options", default=None, help="block size, edges per thread, pack order of the batch")
    ap.add_argument("--duration", type=float, default=DURATION, help="simulated seconds")
    ap.add_argument("--batch-only", action="store_true", help="run the batch alone")
    a = ap.parse_args()
    t0 = time.perf_counter()
    members = members_of(a.config)
    print(f"{a.config}: {len(members)} members generated in {time.perf_counter() - t0:.1f}s", flush=True)
    opts = dict(compensated=True, use_graph=True)
    if a.options:
        bs, ept, order = (int(v) for v in a.options.split(","))
        opts.update(block_size=bs, edges_per_thread=ept, pack_order=order)
    batch = run_hip_ensemble(members, duration_s=a.duration, max_steps_per_hour=BUDGET, **opts)
    batch_s = float(sum(batch[0]["report"]["chunk_runtimes_s"]))
    print(f"batch: {batch_s:.3f}s of advance calls, steps "
          f"{min(r['report']['step_count'] for r in batch)}-{max(r['report']['step_count'] for r in batch)}",
          flush=True)
    if a.batch_only:
        a.out.write_text(json.dumps({"config": a.config, "duration_s": a.duration, "members": len(members),
                                     "batch_advance_seconds": batch_s,
                                     "steps": [int(r["report"]["step_count"]) for r in batch]}) + "\n")
        return
    rows, seq_s, tri_steps = [], 0.0, 0
    for m, (inp, got) in enumerate(zip(members, batch)):
        ref = run_hip_v1_parity(inp, duration_s=a.duration, max_steps_per_hour=BUDGET, keep_final_state=False,
                                compensated=True, use_graph=True)
        seq_s += float(sum(ref["report"]["chunk_runtimes_s"]))
        cells = int(inp.surface.cell_area.shape[0])
        tri_steps += cells * int(got["report"]["step_count"])
        res = gates.validate(ref, got, inp.surface.cell_area)
        row = {"member": m, "triangles": cells, "steps": int(got["report"]["step_count"]),
               "plain_steps": int(ref["report"]["step_count"]),
               "snapshots_bit_identical": bool(np.array_equal(got["depth_snapshots_m"], ref["depth_snapshots_m"])
                                               and np.array_equal(got["drain_node_volume_snapshots_m3"],
                                                                  ref["drain_node_volume_snapshots_m3"])),
               "gates_passed": bool(res["passed"]), "gates": res["gates"], "summary": res["summary"],
               "max_depth_m": float(got["report"]["max_depth_m"]),
               "rainfall_input_m3": float(got["report"]["rainfall_input_m3"]),
               "mass_residual_ratio": float(got["report"]["v1_compatible_mass_residual_ratio"])}
        rows.append(row)
        print(f"member {m}: {cells} triangles, {row['steps']} steps, bit-identical {row['snapshots_bit_identical']}, "
              f"gates {row['gates_passed']}", flush=True)
    out = {"config": a.config, "duration_s": a.duration, "members": len(members), "options": opts,
           "triangles": int(sum(r["triangles"] for r in rows)),
           "batch_advance_seconds": batch_s, "sequential_plain_advance_seconds": seq_s,
           "batch_triangle_steps_per_s": tri_steps / batch_s, "sequential_triangle_steps_per_s": tri_steps / seq_s,
           "speedup_vs_sequential_plain": seq_s / batch_s,
           "all_bit_identical": all(r["snapshots_bit_identical"] for r in rows),
           "all_gates_passed": all(r["gates_passed"] for r in rows), "per_member": rows}
    a.out.write_text(json.dumps(out, indent=1, default=_plain) + "\n")
    print(json.dumps({k: out[k] for k in ("config", "members", "triangles", "batch_advance_seconds",
                                          "sequential_plain_advance_seconds", "speedup_vs_sequential_plain",
                                          "all_bit_identical", "all_gates_passed")}), flush=True)


if __name__ == "__main__":
    main()
