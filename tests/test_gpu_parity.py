"""HIP path vs the CPU oracle on the same seeded inputs, through the C ABI.

Bit-exact equality of every tick's contact-pair set, emitted meetings
(pid_a, pid_b, start, end), active pairs and quadtree leaf boxes."""
import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime
from tests.parity import run_parity

pytestmark = pytest.mark.gpu


def test_device_walk_matches_host_walk():
    w = gen.skewed_density(60_000, 25)
    g = gen.Generator(w)
    dw = runtime.DeviceWalkers(g)
    trace = runtime.DeviceBuffer(16 * w.num_objects * 24)
    dw.walk(24, trace)
    snap = np.zeros(2 * w.num_objects, dtype=np.float64)
    for k in range(24):
        g.step()
        trace.download(snap, k * 16 * w.num_objects)
        assert np.array_equal(snap.view(np.uint64), g.xy.view(np.uint64)), f"tick {k + 1} differs"
    assert np.array_equal(dw.positions().view(np.uint64), g.xy.view(np.uint64))


def test_parity_small_dense_city():
    # 3000 walkers in a 500 m box: dense enough for many contacts and meetings
    w = gen.uniform_city(3000, 60, 500.0, seed=11, grid_capacity=16)
    tot = run_parity(w, 60)
    assert tot["meetings"] > 50 and tot["contacts"] > 1000


def test_parity_min_meet_and_reach_variants():
    for seed, r, mm, cap in [(5, 1.5, 3, 32), (6, 3.0, 10, 64)]:
        w = gen.uniform_city(4000, 40, 600.0, seed=seed, grid_capacity=cap)
        w.reach_distance, w.min_meet_time = r, mm
        run_parity(w, 40)


def test_parity_hotspots_small():
    w = gen.Workload("hot_small", 20_000, 30, 3000.0, seed=21, num_hotspots=8, hotspot_fraction=0.8,
                     hotspot_sigma=100.0, grid_capacity=64)
    tot = run_parity(w, 30)
    assert tot["contacts"] > 10_000


@pytest.mark.parametrize("join_tile,capacity,table_threads",
                         [(256, 600, 512), (512, 1024, 1024), (256, 64, 256), (256, 256, 1024)])
def test_parity_join_tiles_and_leaf_capacities(join_tile, capacity, table_threads):
    """Leaves larger than one LDS tile are joined tile by tile; every tile size,
    leaf capacity and table-update workgroup size gives the oracle's result at
    that capacity."""
    w = gen.skewed_density(40_000, 12)
    w.side = 3000.0
    w.grid_capacity = capacity
    run_parity(w, 12, join_tile=join_tile, table_threads=table_threads, precision=runtime.PRECISION_FP32_GUARD,
               check_leaves_every=6)


def test_parity_plain_cpu_reference_config():
    # 100k objects, 10 km box, r=2, min=5, leaf capacity 200; 100 ticks
    w = gen.plain_cpu_reference()
    tot = run_parity(w, 100, check_contacts=True, check_leaves_every=25)
    assert tot["meetings"] > 100


def test_parity_complete_mode_is_all_pairs():
    w = gen.uniform_city(2500, 20, 400.0, seed=31, grid_capacity=16)
    run_parity(w, 20, complete=True)


def test_parity_fp32_guard_band_equals_fp64():
    # the guard-band form decides every pair outside (r-eps, r+eps] in FP32 and
    # re-tests the rest in FP64: the contact and meeting sets stay bit-identical
    w = gen.Workload("hot_guard", 30_000, 25, 3000.0, seed=41, num_hotspots=8, hotspot_fraction=0.8,
                     hotspot_sigma=80.0, grid_capacity=64)
    run_parity(w, 25, precision=runtime.PRECISION_FP32_GUARD)


@pytest.mark.parametrize("min_meet", [0, 3])
def test_parity_without_recorded_contacts(min_meet):
    """Without record_contacts a contact first seen by the inside join of its
    leaf, and not within reach at the tick before, is not stored: it is
    recognised when it continues from the objects' leaves at the tick before.
    Meetings, distinct-pair counts and the listed pairs still equal the
    oracle's (min_meet_time 0 keeps every contact)."""
    w = gen.Workload("hot_pruned", 30_000, 30, 3000.0, seed=45, num_hotspots=8, hotspot_fraction=0.8,
                     hotspot_sigma=80.0, grid_capacity=64)
    w.min_meet_time = min_meet
    tot = run_parity(w, 30, check_contacts=False, precision=runtime.PRECISION_FP32_GUARD)
    assert tot["meetings"] > 100


def test_guard_band_rechecks_only_inside_band():
    w = gen.uniform_city(5000, 10, 600.0, seed=43, grid_capacity=32)
    g = gen.Generator(w)
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=w.grid_capacity,
                              precision=runtime.PRECISION_FP32_GUARD, guard_epsilon=1e-3)
    with runtime.Workbench(cfg) as wb:
        for t in range(w.ticks):
            xy = g.positions() if t == 0 else g.step()
            wb.tick(xy, t)
        tot = wb.totals()
    # rechecks are pairs whose FP32 distance lies within 1 mm of the 2 m reach
    assert tot["rechecks"] < tot["contacts"]


@pytest.mark.parametrize("graph_ticks", [0, 4])
def test_run_ticks_device_trace_matches_oracle(graph_ticks):
    """Many ticks from a device-resident trace in one call (no host sync per
    tick; with graph_ticks, replays of one captured hipGraph) give the same
    active pairs and meeting totals as the oracle."""
    from oracle.oracle import Oracle
    w = gen.uniform_city(20_000, 40, 1000.0, seed=51, grid_capacity=32)
    g = gen.Generator(w)
    dw = runtime.DeviceWalkers(g)
    trace = runtime.DeviceBuffer(16 * w.num_objects * w.ticks)
    trace.upload(g.positions(), 0)
    dw.walk(w.ticks - 1, trace, 16 * w.num_objects)
    o = Oracle(w.num_objects, w.box(), w.reach_distance, w.grid_capacity, w.min_meet_time)
    om = 0
    for t in range(w.ticks):
        o.tick(g.positions() if t == 0 else g.step(), t)
        om += len(o.meetings())
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=w.grid_capacity, graph_ticks=graph_ticks)
    with runtime.Workbench(cfg) as wb:
        ms = wb.run_ticks(trace.address(0), 0, 15)
        ms2 = wb.run_ticks(trace.address(16 * w.num_objects * 15), 15, w.ticks - 15)
        assert len(ms) == 15 and len(ms2) == w.ticks - 15 and (ms > 0).all()
        tot = wb.totals()
        assert tot["ticks"] == w.ticks and tot["meetings"] == om
        k, s = wb.active_pairs()
        ok, os_ = o.active()
        keep = os_ != w.ticks - 1   # pairs first seen at the last tick are not kept
        assert len(k) > 100 and np.array_equal(k, ok[keep]) and np.array_equal(s, os_[keep])
        assert wb.stats()["num_active"] == len(ok)
        b, lv, _ = wb.leaves()
        ob, olv, _ = o.leaves()
        assert np.array_equal(b, ob) and np.array_equal(lv, olv)
