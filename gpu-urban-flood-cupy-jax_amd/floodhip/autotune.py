"""Autotuner for the batched solver: block size, edges per thread and member
packing order, chosen by timing a short run of the actual batch."""

from __future__ import annotations

import itertools
import time

from .solver import HipEnsemble


def _rate(ens, sizes, steps, horizon):
    before = [c["global_step"] for c, _ in ens.member_clocks()]
    t0 = time.perf_counter()
    ens.advance_to(horizon, horizon, max_steps=steps, raise_on_budget=False)
    el = time.perf_counter() - t0
    after = [c["global_step"] for c, _ in ens.member_clocks()]
    tri = sum(n * (a - b) for n, a, b in zip(sizes, after, before))
    return tri / el, el


def candidates(uneven: bool):
    orders = (0, 1, 2) if uneven else (0,)
    for bs, ept, order in itertools.product((64, 128, 256), (1, 2), orders):
        yield dict(block_size=bs, edges_per_thread=ept, pack_order=order)


def autotune(members, *, steps=64, warmup=16, horizon=3600.0, uneven=None, log=None, **fixed):
    """Returns (best option dict, [(options, triangle-steps/s), ...])."""
    sizes = [int(m.surface.cell_area.shape[0]) for m in members]
    if uneven is None:
        uneven = len(set(sizes)) > 1
    table = []
    for opts in candidates(uneven):
        with HipEnsemble(members, **fixed, **opts) as ens:
            ens.advance_to(horizon, horizon, max_steps=warmup, raise_on_budget=False)
            rate, el = _rate(ens, sizes, steps, horizon)
        table.append((opts, rate))
        if log:
            log(f"autotune {opts}: {rate:.4e} tri-steps/s ({el:.3f}s)")
    best = max(table, key=lambda r: r[1])[0]
    return best, table
