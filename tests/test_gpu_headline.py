"""The path the benchmark times, against the oracle: contacts not recorded
(only the contacts that can continue or end a run are stored), the FP32
guard band, leaves of several join tiles, and ticks replayed from captured
8-tick graphs.

The dense workload (80k objects in four Gaussian hotspots of sigma 20 m)
makes leaves of up to about 1400 objects at capacity 320, as the 10M skewed
workload does: more than two reference leaf stores (2*cap) and up to six
256-object join tiles."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from glint_pipeline_amd import autotune, gen, runtime
from tests.parity import device_trace, run_parity, run_ticks_parity, split_log_by_tick

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).parent / "golden" / "skewed_10m_fixture.json"


def dense(ticks: int, cap: int, seed: int = 61) -> gen.Workload:
    return gen.Workload("dense_80k", 80_000, ticks, 3000.0, seed=seed, num_hotspots=4, hotspot_fraction=0.8,
                        hotspot_sigma=20.0, grid_capacity=cap)


@pytest.mark.parametrize("join_tile,capacity,table_threads",
                         [(256, 320, 512), (512, 320, 1024), (256, 1024, 256), (512, 64, 512)])
def test_pruned_parity_multi_tile_leaves(join_tile, capacity, table_threads):
    """Tick by tick without recorded contacts: every tick's meetings, the
    active-pair count, and the leaf boxes every fourth tick equal the
    oracle's at the same capacity, with leaves of several tiles."""
    w = dense(16, capacity)
    run_parity(w, 16, check_contacts=False, check_leaves_every=4, precision=runtime.PRECISION_FP32_GUARD,
               join_tile=join_tile, table_threads=table_threads, num_meeting_buckets=1 << 21)


def test_pruned_parity_fp64_multi_tile_leaves():
    w = dense(12, 320, seed=62)
    run_parity(w, 12, check_contacts=False, check_leaves_every=4, precision=runtime.PRECISION_FP64,
               join_tile=256, table_threads=512, num_meeting_buckets=1 << 21)


@pytest.mark.parametrize("min_meet", [1, 5])
def test_headline_parameters_graph8_matches_oracle(min_meet):
    """The benchmark's parameters (capacity 320, join tile 256, table
    workgroup 512, FP32 guard band, 8 ticks per graph) on leaves of up to
    ~1400 objects: 5 plain ticks, then two graph replays and a remainder;
    the meeting log equals the oracle's meetings of every tick."""
    w = dense(24, 320)
    w.min_meet_time = min_meet
    res = run_ticks_parity(w, [5, 19], join_tile=256, table_threads=512, graph_ticks=8,
                           num_meeting_buckets=1 << 21, meeting_capacity=1 << 23, contact_capacity=1 << 24,
                           meeting_log_capacity=1 << 26)
    pt = res["per_tick"]
    assert max(s["max_leaf_count"] for s in pt) > 3 * 256, "leaves should span more than three join tiles"
    assert sum(s["ref_drop_leaves"] for s in pt) > 0
    assert res["meetings"] > 10_000


def test_graph_ticks_counts_and_sets_grid_rebuild():
    """The grid rebuild with graphs and the guard band, pruned."""
    w = dense(20, 320, seed=63)
    from glint_pipeline_amd.autotune import box_for
    run_ticks_parity(w, [4, 16], dynamic_schema=runtime.SCHEMA_GRID_REBUILD, box=box_for(w, "grid"),
                     join_tile=256, table_threads=512, graph_ticks=8, num_meeting_buckets=1 << 21)


@pytest.mark.skipif(not FIXTURE.exists(), reason="10M skewed fixture not generated")
def test_skewed_10m_benchmark_configuration_matches_fixture():
    """The benchmark's own configuration at its full size (10M skewed, tuned
    parameters, 8-tick graphs, 5 warm-up ticks then two replays) against the
    oracle's per-tick meeting digests, the reference-divergence counts, the
    active pairs and the leaf boxes written by
    tests/golden/make_headline_golden.py."""
    fx = json.loads(FIXTURE.read_text())
    case = fx["case"]
    w = gen.skewed_density(case["objects"], 300)
    w.grid_capacity = case["grid_capacity"]
    params = autotune.params_for("skewed_10m")
    assert params["grid_capacity"] == case["grid_capacity"]
    T = case["ticks"]
    cfg = autotune.config_for(w, params, autotune.pairs_estimate("skewed", w.num_objects), graph_ticks=8,
                              meeting_log_capacity=80 << 20)
    trace = device_trace(w, T)
    with runtime.Workbench(cfg) as wb:
        wb.run_ticks(trace.address(0), 0, 5, latencies=False)
        wb.run_ticks(trace.address(16 * w.num_objects * 5), 5, T - 5, latencies=False)
        tot = wb.totals()
        ft = fx["ticks"]
        assert tot["distinct_pairs"] == sum(t["contacts"] for t in ft)
        assert tot["meetings"] == sum(t["meetings"] for t in ft)
        assert tot["ref_drop_leaves"] == sum(t["ref_drop_leaves"] for t in ft)
        assert tot["ref_drop_objects"] == sum(t["ref_drop_objects"] for t in ft)
        assert tot["max_leaf_count"] == max(t["max_leaf"] for t in ft)
        if "tests" in ft[0]:
            assert tot["tests"] == sum(t["tests"] for t in ft)
        log = split_log_by_tick(wb.meeting_log(sort=False))
        for t in ft:
            m = log.get(t["tick"], np.zeros((0, 4), dtype=np.uint32))
            assert len(m) == t["meetings"], f"tick {t['tick']}: {len(m)} meetings, oracle {t['meetings']}"
            assert hashlib.sha256(np.ascontiguousarray(m).tobytes()).hexdigest() == t["meetings_sha256"], \
                f"tick {t['tick']}: meeting set differs"
        st = wb.stats()
        assert st["num_active"] == ft[-1]["active"]
        k, s = wb.active_pairs()
        assert len(k) == fx["active_before_last"]
        h = hashlib.sha256(k.tobytes())
        h.update(s.tobytes())
        assert h.hexdigest() == fx["active_before_last_sha256"]
        b, _, _ = wb.leaves()
        assert len(b) == fx["final_leaves"]
        assert hashlib.sha256(np.ascontiguousarray(b).tobytes()).hexdigest() == fx["final_leaves_sha256"]
    trace.free()
