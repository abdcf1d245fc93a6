"""Snapshots of a context at a tick boundary restored into a new context
(glint_snapshot / glint_restore): the restored run continues exactly, so
its meetings of every tick equal the oracle's and the uninterrupted run's,
in graph runs without recorded contacts, tick by tick with recorded
contacts, and with update ingestion and the sorted key-value output."""
import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime
from oracle.oracle import Oracle
from tests.parity import device_trace, split_log_by_tick

pytestmark = pytest.mark.gpu


def oracle_run(w, ticks, box=None):
    g = gen.Generator(w)
    o = Oracle(w.num_objects, box or w.box(), w.reach_distance, w.grid_capacity, w.min_meet_time)
    meetings = {}
    for t in range(ticks):
        o.tick(g.positions() if t == 0 else g.step(), t)
        if len(o.meetings()):
            meetings[t] = o.meetings().copy()
    return o, meetings


def test_restored_graph_run_continues_exactly():
    w = gen.Workload("hot_snap", 40_000, 24, 3000.0, seed=91, num_hotspots=8, hotspot_fraction=0.8,
                     hotspot_sigma=40.0, grid_capacity=64)
    T, cut = 24, 10
    trace = device_trace(w, T)
    o, meetings = oracle_run(w, T)
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=w.grid_capacity, graph_ticks=8,
                              precision=runtime.PRECISION_FP32_GUARD, meeting_log_capacity=1 << 22,
                              num_meeting_buckets=1 << 20)
    with runtime.Workbench(cfg) as a, runtime.Workbench(cfg) as b:
        a.run_ticks(trace.address(0), 0, cut)
        snap = a.snapshot()
        b.restore(snap)
        assert b.stats()["num_active"] == a.stats()["num_active"]
        for wb in (a, b):
            wb.run_ticks(trace.address(16 * w.num_objects * cut), cut, T - cut)
        for wb in (a, b):
            log = split_log_by_tick(wb.meeting_log(sort=False))
            assert sorted(log) == sorted(meetings)
            assert all(np.array_equal(log[t], m) for t, m in meetings.items())
            k, s = wb.active_pairs()
            ok, os_ = o.active()
            keep = os_ != T - 1
            assert np.array_equal(k, ok[keep]) and np.array_equal(s, os_[keep])
            bx, lv, _ = wb.leaves()
            ob, olv, _ = o.leaves()
            assert np.array_equal(bx, ob) and np.array_equal(lv, olv)
        assert sum(len(m) for m in meetings.values()) > 1000
    o.close()
    trace.free()


def test_restored_tick_by_tick_with_recorded_contacts():
    w = gen.uniform_city(8000, 20, 600.0, seed=92, grid_capacity=16)
    g = gen.Generator(w)
    o = Oracle(w.num_objects, w.box(), w.reach_distance, w.grid_capacity, w.min_meet_time)
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=w.grid_capacity, record_contacts=True)
    a = runtime.Workbench(cfg)
    b = None
    for t in range(w.ticks):
        xy = g.positions() if t == 0 else g.step()
        o.tick(xy, t)
        if t == 8:
            b = runtime.Workbench(cfg)
            b.restore(a.snapshot())
            a.close()
            a = b
        m = a.tick(xy, t)
        assert np.array_equal(m, o.meetings()), f"tick {t}: meetings differ"
        assert np.array_equal(a.contacts(), o.contacts()), f"tick {t}: contacts differ"
    a.close()
    o.close()


def test_restored_serving_context_with_key_value_output():
    """Update ingestion (10% moving) continues from the restored positions;
    the key-value flush after the run equals the uninterrupted run's."""
    w = gen.uniform_city(6000, 16, 300.0, seed=93, grid_capacity=16)
    w.move_fraction = 0.1
    N = w.num_objects
    cfg = runtime.make_config(N, w.box(), grid_capacity=w.grid_capacity, meeting_log_capacity=1 << 20,
                              kv_capacity=1 << 22, kv_area_limit=1e30)
    outs = []
    for restore_at in (None, 7):
        g = gen.Generator(w)
        dw = runtime.DeviceWalkers(g)
        pids, upd, cnt = runtime.DeviceBuffer(4 * N), runtime.DeviceBuffer(16 * N), runtime.DeviceBuffer(4)
        full = runtime.DeviceBuffer(16 * N)
        full.upload(g.positions())
        wb = runtime.Workbench(cfg)
        for t in range(w.ticks):
            if t == restore_at:
                nb = runtime.Workbench(cfg)
                nb.restore(wb.snapshot())
                wb.close()
                wb = nb
            if t == 0:
                wb.set_points_device(full.address(), 0)
            else:
                k = dw.step_updates(pids, upd, cnt)
                wb.ingest_updates(pids.address(), upd.address(), k, t)
            wb.filter()
            wb.reachability()
            wb.update_meetings(copy_out=False)
            wb.update_schema()
        fl = wb.kv_flush()
        outs.append((wb.meeting_log(), fl))
        wb.close()
        for buf in (pids, upd, cnt, full):
            buf.free()
    (la, fa), (lb, fb) = outs
    assert len(la) > 50 and np.array_equal(la, lb)
    assert fa["summary"] == fb["summary"]
    for part in ("keys", "boxes", "okeys", "sids", "ctfs", "bitmaps", "packed"):
        assert np.array_equal(fa[part], fb[part]), part


def test_restore_refuses_another_configuration():
    w = gen.uniform_city(3000, 3, 300.0, seed=94, grid_capacity=16)
    g = gen.Generator(w)
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=16)
    with runtime.Workbench(cfg) as a:
        with pytest.raises(runtime.GlintError) as e:
            a.snapshot()   # no tick evaluated yet
        assert e.value.status == runtime.GLINT_ERR_STATE
        a.tick(g.positions(), 0)
        snap = a.snapshot()
    other = runtime.make_config(w.num_objects + 1, w.box(), grid_capacity=16)
    with runtime.Workbench(other) as b:
        with pytest.raises(runtime.GlintError) as e:
            b.restore(snap)
        assert e.value.status == runtime.GLINT_ERR_INVALID_ARGUMENT
