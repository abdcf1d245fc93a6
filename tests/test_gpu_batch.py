"""Batched scenarios: cities of uneven size, reach and minimum duration
ticked together (one context per city, one fused launch per step for all of
them, optionally one graph) give, city by city, exactly the meetings of the
same city run alone in the plain per-tick form, and the oracle's for the
smallest cities; cities with different kernel variants go to separate
groups, and more cities than a group holds to more groups."""
import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime

pytestmark = pytest.mark.gpu

TICKS = 26


def cities():
    return gen.batched_cities(count=6, ticks=TICKS, seed=9, n_min=2_000, n_max=30_000, grid_capacity=32)


def traces(ws):
    out = []
    for w in ws:
        g = gen.Generator(w)
        tr = np.zeros((TICKS, w.num_objects, 2))
        for t in range(TICKS):
            tr[t] = (g.positions() if t == 0 else g.step()).reshape(-1, 2)
        out.append(tr)
    return out


def cfg_of(w, log_cap=1 << 20, precision=runtime.PRECISION_FP64):
    return runtime.make_config(w.num_objects, w.box(), reach_distance=w.reach_distance,
                               grid_capacity=w.grid_capacity, min_meet_time=w.min_meet_time,
                               meeting_log_capacity=log_cap, precision=precision)


def alone(w, tr):
    """The plain form: host positions, one call sequence per tick."""
    with runtime.Workbench(cfg_of(w)) as wb:
        for t in range(TICKS):
            wb.tick(tr[t].reshape(-1), t)
        return wb.meeting_log(), wb.active_pairs()


@pytest.mark.parametrize("graph_ticks", [0, 4])
def test_batch_matches_each_city_alone(graph_ticks):
    ws = cities()
    trs = traces(ws)
    print(f"[batch] graph_ticks={graph_ticks} sizes={[w.num_objects for w in ws]}", flush=True)
    bufs = []
    for tr in trs:
        b = runtime.DeviceBuffer(tr.nbytes)
        b.upload(tr)
        bufs.append(b)
    with runtime.Batch([cfg_of(w) for w in ws], graph_ticks=graph_ticks) as bt:
        # two calls: a remainder after whole graphs, then graphs replayed again
        bt.run_ticks([b.address(0) for b in bufs], 0, 10)
        ms = bt.run_ticks([b.address(10 * w.num_objects * 16) for b, w in zip(bufs, ws)], 10, TICKS - 10)
        assert (ms > 0).all()
        for i, (w, tr) in enumerate(zip(ws, trs)):
            got_log = bt.contexts[i].meeting_log()
            got_act = bt.contexts[i].active_pairs()
            want_log, want_act = alone(w, tr)
            assert len(want_log) > 0, w.name
            assert np.array_equal(got_log, want_log), f"{w.name}: meetings"
            for a, b in zip(got_act, want_act):
                assert np.array_equal(a, b), f"{w.name}: active pairs"
    for b in bufs:
        b.free()


def test_batch_smallest_cities_match_oracle():
    from oracle.oracle import Oracle
    ws = sorted(cities(), key=lambda w: w.num_objects)[:2]
    trs = traces(ws)
    bufs = []
    for tr in trs:
        b = runtime.DeviceBuffer(tr.nbytes)
        b.upload(tr)
        bufs.append(b)
    with runtime.Batch([cfg_of(w, precision=runtime.PRECISION_FP32_GUARD) for w in ws], graph_ticks=8) as bt:
        bt.run_ticks([b.address(0) for b in bufs], 0, TICKS)
        for i, (w, tr) in enumerate(zip(ws, trs)):
            o = Oracle(w.num_objects, w.box(), w.reach_distance, w.grid_capacity, w.min_meet_time)
            want = []
            for t in range(TICKS):
                o.tick(tr[t].reshape(-1), t)
                want.append(o.meetings())
            o.close()
            want = np.concatenate(want)
            want = want[np.lexsort((want[:, 3], want[:, 2], want[:, 1], want[:, 0]))]
            assert np.array_equal(bt.contexts[i].meeting_log(), want), w.name
    for b in bufs:
        b.free()


def test_meeting_log_ring_keeps_the_newest():
    w = gen.uniform_city(4000, TICKS, 600.0, seed=3, grid_capacity=32)
    tr = traces([w])[0]
    full, _ = alone(w, tr)
    with runtime.Workbench(cfg_of(w, log_cap=64)) as wb:
        per_tick = []
        for t in range(TICKS):
            per_tick.append(wb.tick(tr[t].reshape(-1), t))
        total, first = wb.meeting_log_info()
        assert total == len(full) and first == max(0, total - 64)
        got = wb.meeting_log()
        emitted = np.concatenate([m for m in per_tick if len(m)])
        assert len(got) == min(64, total)
        # the retained entries are the newest ones: every tick after the
        # oldest retained one is complete
        last_ticks = {int(m[3]) + 1 for m in got}
        newest = max(last_ticks)
        assert newest == max(int(m[3]) + 1 for m in emitted)
        assert first > 0
        with pytest.raises(runtime.GlintError):
            wb.meeting_log(start=0)


def variant_cfg(w, v):
    """City w with setting v: the kernel variants a batch fuses separately
    (join tile, table workgroup, key-value output, schema mode) and the
    settings fused launches share (precision, meeting log on or off)."""
    kw = dict(reach_distance=w.reach_distance, grid_capacity=w.grid_capacity, min_meet_time=w.min_meet_time,
              meeting_log_capacity=1 << 20)
    box = w.box()
    if v == 1:
        kw["join_tile"] = 512
    elif v == 2:
        kw["dynamic_schema"] = runtime.SCHEMA_GRID_REBUILD
        box = (-1.0, -1.0, w.side + 1.0, w.side + 1.0)   # no positions on the grid's high y edge
    elif v == 3:
        kw["dynamic_schema"] = runtime.SCHEMA_STATIC
    elif v == 4:
        kw["table_threads"] = 256
    elif v == 5:
        kw.update(table_threads=1024, precision=runtime.PRECISION_FP32_GUARD)
    elif v == 6:
        kw.update(kv_capacity=1 << 21, kv_area_limit=1e30)
    elif v == 7:
        kw.update(precision=runtime.PRECISION_FP32_GUARD, meeting_log_capacity=0)
    return runtime.make_config(w.num_objects, box, **kw)


def run_alone(cfg, buf, ticks):
    """The same city in a context of its own, plain launches from the same trace."""
    with runtime.Workbench(cfg) as wb:
        wb.run_ticks(buf.address(0), 0, ticks)
        return outcome(wb, cfg)


def outcome(wb, cfg):
    out = {"active": wb.active_pairs(), "leaves": wb.leaves(), "totals": wb.totals()}
    if cfg.meeting_log_capacity:
        out["log"] = wb.meeting_log()
    if cfg.kv_capacity:
        out["kv"] = wb.kv_flush()
    return out


def assert_same(got, want, where):
    for part in ("active", "leaves"):
        for a, b in zip(got[part], want[part]):
            assert np.array_equal(a, b), f"{where}: {part}"
    for k in ("meetings", "distinct_pairs", "tests", "ref_drop_objects", "max_leaf_count"):
        assert got["totals"][k] == want["totals"][k], f"{where}: {k}"
    if "log" in want:
        assert np.array_equal(got["log"], want["log"]), f"{where}: meetings"
    if "kv" in want:
        assert got["kv"]["summary"] == want["kv"]["summary"], where
        for part in ("keys", "boxes", "sids", "ctfs", "bitmaps", "packed"):
            assert np.array_equal(got["kv"][part], want["kv"][part]), f"{where}: kv {part}"


@pytest.mark.parametrize("graph_ticks", [0, 4])
def test_mixed_settings_in_fused_groups(graph_ticks):
    """Ten cities with eight settings in seven groups of fused launches
    (settings 0, 7 and 8 share one, 1 and 9 another), each city equal to
    itself run alone."""
    ws = gen.batched_cities(count=10, ticks=TICKS, seed=21, n_min=2_000, n_max=20_000, grid_capacity=32)
    cfgs = [variant_cfg(w, i % 8) for i, w in enumerate(ws)]
    trs = traces(ws)
    bufs = []
    for tr in trs:
        b = runtime.DeviceBuffer(tr.nbytes)
        b.upload(tr)
        bufs.append(b)
    print(f"[batch] mixed settings, graph_ticks={graph_ticks}", flush=True)
    with runtime.Batch(cfgs, graph_ticks=graph_ticks) as bt:
        assert bt.groups() == 7
        bt.run_ticks([b.address(0) for b in bufs], 0, 9)
        bt.run_ticks([b.address(9 * w.num_objects * 16) for b, w in zip(bufs, ws)], 9, TICKS - 9)
        got = [outcome(c, cfg) for c, cfg in zip(bt.contexts, cfgs)]
    for i, (cfg, buf) in enumerate(zip(cfgs, bufs)):
        want = run_alone(cfg, buf, TICKS)
        assert want["totals"]["meetings"] > 0, ws[i].name
        assert_same(got[i], want, f"{ws[i].name} (setting {i % 8})")
    for b in bufs:
        b.free()


def test_more_cities_than_one_group_holds():
    """70 alike cities: groups of 64 and 6."""
    ws = gen.batched_cities(count=70, ticks=12, seed=23, n_min=500, n_max=3_000, grid_capacity=16)
    cfgs = [runtime.make_config(w.num_objects, w.box(), reach_distance=w.reach_distance, grid_capacity=16,
                                min_meet_time=w.min_meet_time, meeting_log_capacity=1 << 18,
                                num_meeting_buckets=1 << 14) for w in ws]
    bufs = []
    for w in ws:
        g = gen.Generator(w)
        tr = np.stack([g.positions().copy()] + [g.step().copy() for _ in range(11)])
        b = runtime.DeviceBuffer(tr.nbytes)
        b.upload(tr)
        bufs.append(b)
    print("[batch] 70 cities", flush=True)
    with runtime.Batch(cfgs, graph_ticks=4) as bt:
        assert bt.groups() == 2
        bt.run_ticks([b.address(0) for b in bufs], 0, 12)
        got = [outcome(c, cfg) for c, cfg in zip(bt.contexts, cfgs)]
    meetings = 0
    for i, (cfg, buf) in enumerate(zip(cfgs, bufs)):
        want = run_alone(cfg, buf, 12)
        meetings += want["totals"]["meetings"]
        assert_same(got[i], want, ws[i].name)
    assert meetings > 0
    for b in bufs:
        b.free()
