"""HIP path vs the CPU oracle on the same seeded inputs, through the C ABI.

Bit-exact equality of every tick's contact-pair set, emitted meetings
(pid_a, pid_b, start, end), active pairs and quadtree leaf boxes."""
import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime
from tests.parity import run_parity, run_ticks_parity

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


def test_guard_band_rechecks_exactly_the_band():
    """One leaf (N below capacity) so every pair is tested once, inside its
    leaf.  Pairs planted at r + delta, |delta| <= 2 eps, and random ones.
    The FP32 distance of a pair differs from its FP64 distance by less than
    eps/2 on leaves this small, so every pair within eps/2 of r is re-tested
    in FP64 and none farther than 3 eps/2 from it: the GPU's recheck count
    lies between the two brute-force counts.  The contacts equal the
    oracle's."""
    from oracle.oracle import Oracle
    rng = np.random.default_rng(47)
    r, eps, side = 2.0, 1e-3, 100.0
    npairs, nrand = 1500, 1000
    a = rng.uniform(5.0, side - 5.0, (npairs, 2))
    th = rng.uniform(0.0, 2 * np.pi, npairs)
    d = r + rng.uniform(-2 * eps, 2 * eps, npairs)
    b = a + np.stack([d * np.cos(th), d * np.sin(th)], axis=1)
    xy = np.concatenate([a, b, rng.uniform(0.0, side, (nrand, 2))])
    n = len(xy)
    dx = xy[:, None, 0] - xy[None, :, 0]
    dy = xy[:, None, 1] - xy[None, :, 1]
    dist = np.sqrt(dx * dx + dy * dy)[np.triu_indices(n, 1)]
    must = int(np.count_nonzero(np.abs(dist - r) <= eps / 2))
    may = int(np.count_nonzero(np.abs(dist - r) < 1.5 * eps))
    assert must > 300
    box = (0.0, 0.0, side, side)
    o = Oracle(n, box, r, 10 * n, 5)
    o.tick(xy.reshape(-1), 0)
    cfg = runtime.make_config(n, box, reach_distance=r, grid_capacity=10 * n, record_contacts=True,
                              precision=runtime.PRECISION_FP32_GUARD, guard_epsilon=eps)
    with runtime.Workbench(cfg) as wb:
        wb.tick(xy.reshape(-1), 0)
        st = wb.stats()
        assert st["num_leaves"] == 1
        assert must <= st["num_rechecks"] <= may, (must, st["num_rechecks"], may)
        assert np.array_equal(wb.contacts(), o.contacts())
    o.close()


@pytest.mark.parametrize("graph_ticks", [0, 4])
def test_run_ticks_device_trace_matches_oracle(graph_ticks):
    """Many ticks from a device-resident trace in one call (no host sync per
    tick; with graph_ticks, replays of one captured hipGraph): every tick's
    meetings (through the meeting log), the active pairs and the leaves equal
    the oracle's."""
    w = gen.uniform_city(20_000, 40, 1000.0, seed=51, grid_capacity=32)
    res = run_ticks_parity(w, [15, 25], precision=runtime.PRECISION_FP64, graph_ticks=graph_ticks)
    assert res["meetings"] > 100
