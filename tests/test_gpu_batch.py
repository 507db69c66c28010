"""Batched ensemble: members advance together (each with its own CFL dt, clock,
forcing and ledger) and every member must equal a plain single-scenario run of
the same inputs bit for bit -- the kernels do the same arithmetic per cell and
edge, and the per-member block partials line up with the plain run's blocks."""

import numpy as np
import pytest

from floodhip import synth
from floodhip.solver import HipEnsemble, HipParitySolver
from helpers import small_network_scenario, strip_network

pytestmark = pytest.mark.gpu


def _members():
    a = synth.make_plain_small(n=30, extent=1200.0)
    h0 = np.maximum(0.0, 1.5 - a.surface.bed).astype(np.float32)
    a = a._replace(initial_state=a.initial_state._replace(h=h0))
    b = small_network_scenario(seed=5, h_init=0.3)
    c = synth.make_catchment(6000, seed=synth.BASE_SEED + 77, rain_hours=1, rain_mm_h=[90.0])
    d = strip_network(small_network_scenario(seed=9, h_init=0.3), links=True, inlets=False)   # nodes, no pipes
    return [a, b, c, d]


def _plain(inp, **kw):
    with HipParitySolver(inp.surface, inp.coupling, inp.pipe, inp.recharge, inp.forcing, inp.initial_state,
                         **kw) as s:
        clock, ledger = s.advance_to(45.0, 600.0, max_steps=100_000)
        return s.get_state(), clock, ledger


@pytest.mark.parametrize("opts", [dict(), dict(pack_order=1, edges_per_thread=2),
                                  dict(pack_order=2, edges_per_thread=4, block_size=128)])
def test_members_equal_plain_runs_bitwise(opts):
    members = _members()
    ref = [_plain(m) for m in members]
    with HipEnsemble(members, steps_per_launch=16, **opts) as ens:
        ens.advance_to(45.0, 600.0, max_steps=100_000)
        states = ens.member_states()
        clocks = ens.member_clocks()
    for m, (st, (clock, ledger)) in enumerate(zip(states, clocks)):
        rst, rclock, rledger = ref[m]
        assert clock["global_step"] == rclock["global_step"], m
        assert clock["time_s"] == rclock["time_s"], m
        for name in ("h", "hu", "hv", "wet_time_s", "node_volume", "link_flow", "recharge_storage"):
            assert np.array_equal(getattr(st, name), getattr(rst, name)), (m, name)
        assert np.allclose(np.asarray(ledger), np.asarray(rledger), rtol=1e-12, atol=1e-12), m


def test_members_keep_their_own_dt():
    # a deep dam break has a CFL dt well below the 1 s cap the shallow members sit at
    deep = synth.make_dam_break(nx=200, ny=4, length=1000.0, width=20.0, h_up=6.0, h_down=0.5)
    members = _members() + [deep]
    with HipEnsemble(members) as ens:
        ens.advance_to(20.0, 600.0)
        steps = [c["global_step"] for c, _ in ens.member_clocks()]
    assert steps[0] == 20 and steps[-1] > 20   # the deep member takes more (smaller) steps


def _shared_members(n=3):
    # one mesh and network (with pipes, inlets and outfalls), per-member rain and Horton
    base = small_network_scenario(seed=11, n=40, h_init=0.2)
    return synth.make_shared_mesh_ensemble(base, n_members=n, rain_range_mm=(60.0, 240.0))


@pytest.mark.parametrize("opts", [dict(), dict(compensated=True, edges_per_thread=2),
                                  dict(edges_per_thread=4, block_size=128, cell_order=0),
                                  dict(surface_mode="local_inertial"),
                                  dict(surface_mode="local_inertial", compensated=True, block_size=64)])
def test_shared_mesh_members_equal_plain_runs_bitwise(opts):
    """A batch of one mesh stores the edge and cell->edge tables once (member-local
    indices, the members' copies of a tile run back to back); each member must
    still equal its plain run bit for bit, and the copy-per-member batch too."""
    members = _shared_members()
    ref = [_plain(m, **opts) for m in members]
    bytes_by_mode = {}
    for share in (True, False):
        with HipEnsemble(members, steps_per_launch=16, share_mesh=share, **opts) as ens:
            ens.advance_to(45.0, 600.0, max_steps=100_000)
            states = ens.member_states()
            clocks = ens.member_clocks()
            bytes_by_mode[share] = ens.device_bytes()
        for m, (st, (clock, ledger)) in enumerate(zip(states, clocks)):
            rst, rclock, rledger = ref[m]
            assert clock["global_step"] == rclock["global_step"], (share, m)
            assert clock["time_s"] == rclock["time_s"], (share, m)
            for name in ("h", "hu", "hv", "wet_time_s", "node_volume", "link_flow"):
                assert np.array_equal(getattr(st, name), getattr(rst, name)), (share, m, name)
            assert np.allclose(np.asarray(ledger), np.asarray(rledger), rtol=1e-12, atol=1e-12), (share, m)
    # the shared batch holds one copy of the 20 B per edge and 12 B per cell tables
    # (the device keeps at most every input edge: 90 % is a safe lower bound)
    s = members[0].surface
    saved = (len(members) - 1) * (20 * s.edge_left.size + 12 * s.bed.size)
    assert bytes_by_mode[False] - bytes_by_mode[True] >= 0.9 * saved, bytes_by_mode


def test_shared_mesh_batch_volume_against_fp64():
    """The compensated float32 batch against the same batch in FP64, same fixed dt:
    every member's stored volume (surface + nodes) within 1e-5 relative."""
    from floodhip.precision import ensemble_vs_fp64
    res = ensemble_vs_fp64(_shared_members(4), steps=200, dt=0.5)
    worst = res["worst"]
    assert worst["stored_volume_rel"] <= 1e-5, res
    assert worst["surface_volume_rel"] <= 1e-5, res
    assert worst["depth_linf_m"] <= 1e-3, res
    assert all(m["steps"] == 200 for m in res["members"])


def test_same_geometry_other_active_cells_is_not_shared():
    """Members with the same triangles but another active-cell mask have other
    cell->edge tables: the batch keeps a copy per member (same footprint as with
    sharing off) and every member still equals its plain run."""
    a, b = _shared_members(2)
    act = np.asarray(b.surface.active).copy()
    act[::7] = False
    b = b._replace(surface=b.surface._replace(active=act))
    members = [a, b]
    ref = [_plain(m) for m in members]
    footprint = {}
    for share in (True, False):
        with HipEnsemble(members, steps_per_launch=16, share_mesh=share) as ens:
            ens.advance_to(45.0, 600.0, max_steps=100_000)
            states = ens.member_states()
            footprint[share] = ens.device_bytes()
        for m, st in enumerate(states):
            for name in ("h", "hu", "hv", "node_volume"):
                assert np.array_equal(getattr(st, name), getattr(ref[m][0], name)), (share, m, name)
    assert footprint[True] == footprint[False], footprint


def test_ensemble_runner_matches_plain_runner_per_member():
    """runner.run_hip_ensemble (one batch to the output times) against
    runner.run_hip_v1_parity on each member alone: the snapshots are bit-identical,
    the step counts and ledgers equal, and the eight gates pass with zero error."""
    from floodhip import gates
    from floodhip.runner import run_hip_ensemble, run_hip_v1_parity
    from floodhip.solver import BATCH_BLOCK_SIZE
    members = _members()
    batch = run_hip_ensemble(members, duration_s=90.0, snapshot_interval_s=30.0, compensated=True)
    assert len(batch) == len(members)
    for inp, got in zip(members, batch):
        # the batch's default launch shape, so that the FP64 ledger adds the same
        # block partials in the same order and the ledgers are equal, not just close
        ref = run_hip_v1_parity(inp, duration_s=90.0, snapshot_interval_s=30.0, compensated=True,
                                block_size=BATCH_BLOCK_SIZE)
        assert np.array_equal(got["snapshot_times_s"], ref["snapshot_times_s"])
        assert np.array_equal(got["depth_snapshots_m"], ref["depth_snapshots_m"])
        assert np.array_equal(got["drain_node_volume_snapshots_m3"], ref["drain_node_volume_snapshots_m3"])
        assert got["report"]["step_count"] == ref["report"]["step_count"]
        assert got["report"]["steps_at_snapshot"] == ref["report"]["steps_at_snapshot"]
        for k in ("rainfall_input_m3", "infiltration_loss_m3", "outfall_discharge_m3", "surface_capture_m3"):
            assert got["report"][k] == ref["report"][k], k
        res = gates.validate(ref, got, inp.surface.cell_area)
        # the gates' CSI is intersection / max(union, 1): 0 on a snapshot where
        # neither run has a cell at 5 cm or more (the dry-start member's first
        # snapshot), 1 on every other snapshot of identical runs
        for hour in res["hourly"]:
            wet = bool((ref["depth_snapshots_m"][hour["snapshot_index"]] >= gates.DEPTH_THRESHOLD_M).any())
            assert hour["flood_extent_csi"] == (1.0 if wet else 0.0)
        failed = {k for k, ok in res["gates"].items() if not ok}
        assert failed <= {"all_hour_csi_ge_0_95"}, res["gates"]
        assert res["summary"]["maximum_hourly_valid_depth_rmse_m"] == 0.0
