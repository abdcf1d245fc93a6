"""Failure paths on the device, forced from the host: every one returns a
status code (the reference exits or asserts, cuda_util.cuh:19-27)."""
import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime
from tests.parity import run_parity

pytestmark = pytest.mark.gpu


def small():
    w = gen.uniform_city(2000, 5, 400.0, seed=61, grid_capacity=16)
    return w, gen.Generator(w)


def test_nan_and_out_of_box_positions():
    w, g = small()
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=16)
    for bad in (np.nan, -1.0, w.side + 0.5):
        xy = g.positions().copy()
        xy[10] = bad
        with runtime.Workbench(cfg) as wb:
            wb.set_points(xy, 0)
            with pytest.raises(runtime.GlintError) as e:
                wb.filter()
                wb.reachability()
                wb.update_meetings()
            assert e.value.status == 4


def test_call_order_is_checked():
    w, g = small()
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=16)
    with runtime.Workbench(cfg) as wb:
        with pytest.raises(runtime.GlintError) as e:
            wb.reachability()
        assert e.value.status == 9
        wb.set_points(g.positions(), 5)
        wb.filter()
        wb.reachability()
        wb.update_meetings()
        wb.update_schema()
        with pytest.raises(runtime.GlintError) as e:
            wb.set_points(g.positions(), 5)      # ticks must increase
        assert e.value.status == 9


@pytest.mark.parametrize("stage", ["partition", "refine", "evaluate", "update_schema"])
def test_injected_failures(stage):
    w, g = small()
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=16)
    with runtime.Workbench(cfg) as wb:
        wb.inject_failure(stage)
        with pytest.raises(runtime.GlintError) as e:
            wb.tick(g.positions(), 0)
        assert e.value.status == 10


def test_table_grows_past_the_configured_slots():
    """A table configured with one region of 2048 slots: from tick 1 on the
    contacts that may continue a run outnumber its slots, so the ticks use
    more regions (the shape is chosen on the device from each tick's stored
    contacts) and the results still equal the oracle's."""
    w = gen.uniform_city(6000, 12, 200.0, seed=62, grid_capacity=16)
    run_parity(w, 12, check_contacts=False, precision=runtime.PRECISION_FP32_GUARD, num_meeting_buckets=64,
               check_leaves_every=4)
    g = gen.Generator(w)
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=16, num_meeting_buckets=64)
    with runtime.Workbench(cfg) as wb:
        for t in range(3):
            wb.tick(g.positions() if t == 0 else g.step(), t)
        st = wb.stats()
        assert st["num_stored"] > 2048 and st["table_slots"] > 2048


def test_table_shape_follows_density_both_ways():
    """A dense cluster with small moves every tick; at ticks 6..10 half of
    the objects scatter over the box and at 11 they return.  The table's
    region count falls and rises with the stored contacts, and pairs that
    continue across a change are found in old regions of another shape:
    meetings, counts and active pairs still equal the oracle's."""
    from tests.parity import run_positions_parity
    rng = np.random.default_rng(65)
    n, side = 20_000, 1000.0
    home = rng.uniform(470.0, 530.0, (n, 2))
    away = rng.uniform(0.0, side, (n, 2))
    xys = []
    for t in range(16):
        xy = home + rng.uniform(-0.3, 0.3, (n, 2))
        if 6 <= t <= 10:
            xy[1::2] = away[1::2] + rng.uniform(-0.3, 0.3, (n // 2, 2))
        xys.append(np.ascontiguousarray(np.clip(xy, 0.0, side)).reshape(-1))
    st = run_positions_parity(xys, (0.0, 0.0, side, side), grid_capacity=64,
                              precision=runtime.PRECISION_FP32_GUARD, num_meeting_buckets=2048,
                              contact_capacity=1 << 24, meeting_capacity=1 << 22)
    slots = [s["table_slots"] for s in st]
    assert max(slots[1:6]) > slots[8] < max(slots[12:]), slots


def test_unit_capacity_overflow_is_reported():
    w = gen.uniform_city(3000, 3, 300.0, seed=63, grid_capacity=8)
    g = gen.Generator(w)
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=8, unit_capacity=16)
    with runtime.Workbench(cfg) as wb:
        with pytest.raises(runtime.GlintError) as e:
            wb.tick(g.positions(), 0)
        assert e.value.status == 7


def test_node_pool_exhaustion_is_reported():
    w = gen.uniform_city(20000, 3, 2000.0, seed=64, grid_capacity=4)
    g = gen.Generator(w)
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=4, node_capacity=64)
    with runtime.Workbench(cfg) as wb:
        with pytest.raises(runtime.GlintError) as e:
            wb.tick(g.positions(), 0)
        assert e.value.status == 6
