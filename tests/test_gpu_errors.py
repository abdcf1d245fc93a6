"""Failure paths on the device, forced from the host: every one returns a
status code (the reference exits or asserts, cuda_util.cuh:19-27)."""
import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime

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


def test_table_too_small_reports_table_full():
    # one region of 2048 slots; from tick 1 on, the contacts that may continue
    # a run (within reach at the tick before) outnumber its slots
    w = gen.uniform_city(6000, 4, 200.0, seed=62, grid_capacity=16)
    g = gen.Generator(w)
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=16, num_meeting_buckets=64)
    with runtime.Workbench(cfg) as wb:
        with pytest.raises(runtime.GlintError) as e:
            for t in range(w.ticks):
                wb.tick(g.positions() if t == 0 else g.step(), t)
        assert e.value.status == 5


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
