"""Batched scenarios: cities of uneven size, reach and minimum duration
ticked together (one context per city, optionally one graph for all of
them) give, city by city, exactly the meetings of the same city run alone in
the plain per-tick form, and the oracle's for the smallest cities."""
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
