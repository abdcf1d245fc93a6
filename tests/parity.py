"""Shared driver of the GPU-vs-oracle parity tests: the same seeded positions
go to the HIP path (through the C ABI) and to the CPU oracle, tick by tick."""
from __future__ import annotations

import numpy as np

from glint_pipeline_amd import gen, runtime
from oracle.oracle import Oracle


def keys_of(meetings: np.ndarray) -> np.ndarray:
    return (meetings[:, 0].astype(np.uint64) << np.uint64(32)) | meetings[:, 1].astype(np.uint64)


def run_parity(w: gen.Workload, ticks: int, *, check_contacts: bool = True, check_leaves_every: int = 10,
               complete: bool = False, precision: int = runtime.PRECISION_FP64,
               dynamic_schema: int = runtime.SCHEMA_DYNAMIC, box=None, cfg=None, positions=None,
               geography: bool = False, **cfg_kw) -> dict:
    """Runs both paths for `ticks` ticks and asserts identical per-tick contact
    sets, meeting sets and leaf boxes (cfg: a complete configuration to use
    instead of one made from the arguments; positions(t): tick t's positions
    instead of the workload's walk).  Returns totals for the caller."""
    g = gen.Generator(w)
    box = box or w.box()
    o = Oracle(w.num_objects, box, w.reach_distance, w.grid_capacity, w.min_meet_time, complete=complete,
               dynamic_schema=dynamic_schema, geography=geography)
    kw = dict(reach_distance=w.reach_distance, grid_capacity=w.grid_capacity, min_meet_time=w.min_meet_time,
              complete=complete, precision=precision, record_contacts=check_contacts,
              dynamic_schema=dynamic_schema, geography=geography)
    kw.update(cfg_kw)
    if cfg is None:
        cfg = runtime.make_config(w.num_objects, box, **kw)
    check_contacts = check_contacts and bool(cfg.record_contacts)
    totals = {"meetings": 0, "contacts": 0}
    with runtime.Workbench(cfg) as wb:
        for t in range(ticks):
            xy = positions(t) if positions else g.positions() if t == 0 else g.step()
            o.tick(xy, t)
            m = wb.tick(xy, t)
            om = o.meetings()
            assert m.shape == om.shape and np.array_equal(m, om), (
                f"tick {t}: meetings differ ({len(m)} GPU vs {len(om)} oracle)")
            totals["meetings"] += len(m)
            if check_contacts:
                c = wb.contacts()
                oc = o.contacts()
                assert np.array_equal(c, oc), f"tick {t}: contacts differ ({len(c)} GPU vs {len(oc)} oracle)"
                totals["contacts"] += len(c)
            st = wb.stats()
            ost = o.stats()
            assert st["num_active"] == ost["num_active"], f"tick {t}: active pair counts differ"
            assert st["num_tests"] == ost["num_tests"], f"tick {t}: tested pairs differ"
            assert st["ref_drop_objects"] == ost["ref_drop_objects"], f"tick {t}: reference drop counts differ"
            if check_leaves_every and t % check_leaves_every == 0:
                b, lv, cnt = wb.leaves()
                ob, olv, ocnt = o.leaves()
                assert np.array_equal(b, ob) and np.array_equal(lv, olv), f"tick {t}: leaf boxes differ"
        k, s = wb.active_pairs()
        ok, os_ = o.active()
        if not (cfg.record_contacts or cfg.geography) and w.min_meet_time != 0:
            # pairs first seen at the last tick are not kept without recorded contacts
            keep = os_ != ticks - 1
            ok, os_ = ok[keep], os_[keep]
        assert np.array_equal(k, ok) and np.array_equal(s, os_), "active pairs after the last tick differ"
    o.close()
    return totals


def split_log_by_tick(log: np.ndarray) -> dict:
    """Meeting-log rows grouped by the tick that emitted them (end + 1), each
    group sorted by (pid_a, pid_b, start).  The log holds one tick's
    meetings after another, so `end` never decreases along it."""
    if len(log) == 0:
        return {}
    end = log[:, 3]
    assert (np.diff(end.astype(np.int64)) >= 0).all(), "meeting log is not in emission order"
    cuts = np.flatnonzero(np.diff(end)) + 1
    out = {}
    for part in np.split(log, cuts):
        order = np.lexsort((part[:, 2], part[:, 1], part[:, 0]))
        out[int(part[0, 3]) + 1] = part[order]
    return out


def device_trace(w: gen.Workload, ticks: int):
    """Positions of ticks 0..ticks-1 of the workload in device memory
    (tick-major FP64 AoS), walked on the device from the host tick 0."""
    g = gen.Generator(w)
    trace = runtime.DeviceBuffer(16 * w.num_objects * ticks)
    trace.upload(g.positions(), 0)
    if ticks > 1:
        runtime.DeviceWalkers(g).walk(ticks - 1, trace, 16 * w.num_objects)
    return trace


def run_ticks_parity(w: gen.Workload, runs: list[int], *, precision: int = runtime.PRECISION_FP32_GUARD,
                     dynamic_schema: int = runtime.SCHEMA_DYNAMIC, box=None, trace=None, extra=None,
                     geography: bool = False, **cfg_kw) -> dict:
    """Ticks through glint_run_ticks in calls of `runs` ticks each (a call of
    at least graph_ticks ticks replays the captured graph), without recorded
    contacts unless asked: the path the benchmark times.  Asserts that the
    meeting log equals the oracle's meetings of every tick element by element,
    that the distinct contacts, meetings and reference-divergence counts per
    run equal the oracle's sums, and that the active pairs and leaf boxes
    after the last tick are equal; extra(wb, oracle), if given, checks more.
    The positions are the workload's walk, or `trace` (device memory, with
    host_ticks(t) giving tick t's positions for the oracle).  Returns the
    oracle's per-tick statistics."""
    ticks = sum(runs)
    box = box or w.box()
    own = trace is None
    if own:
        trace = device_trace(w, ticks)
        g = gen.Generator(w)
        host_tick = lambda t: g.positions() if t == 0 else g.step()   # noqa: E731
    else:
        trace, host_tick = trace
    o = Oracle(w.num_objects, box, w.reach_distance, w.grid_capacity, w.min_meet_time,
               dynamic_schema=dynamic_schema, geography=geography)
    per_tick, meetings = [], {}
    for t in range(ticks):
        o.tick(host_tick(t), t)
        st = o.stats()
        per_tick.append(st)
        m = o.meetings()
        if len(m):
            meetings[t] = m.copy()
    kw = dict(reach_distance=w.reach_distance, grid_capacity=w.grid_capacity, min_meet_time=w.min_meet_time,
              precision=precision, dynamic_schema=dynamic_schema, meeting_log_capacity=1 << 22,
              geography=geography)
    kw.update(cfg_kw)
    cfg = runtime.make_config(w.num_objects, box, **kw)
    with runtime.Workbench(cfg) as wb:
        first = 0
        for n in runs:
            wb.totals(reset=True)
            ms = wb.run_ticks(trace.address(16 * w.num_objects * first), first, n)
            assert len(ms) == n and (ms > 0).all()
            tot = wb.totals()
            sl = per_tick[first:first + n]
            assert tot["ticks"] == n
            assert tot["distinct_pairs"] == sum(s["num_contacts"] for s in sl), f"ticks {first}..: distinct pairs"
            assert tot["meetings"] == sum(s["num_meetings"] for s in sl), f"ticks {first}..: meeting counts"
            assert tot["ref_drop_leaves"] == sum(s["ref_drop_leaves"] for s in sl)
            assert tot["ref_drop_objects"] == sum(s["ref_drop_objects"] for s in sl)
            assert tot["max_leaf_count"] == max(s["max_leaf_count"] for s in sl)
            assert tot["tests"] == sum(s["num_tests"] for s in sl), f"ticks {first}..: tested pairs"
            first += n
        assert wb.meeting_log_info()[1] == 0, "the meeting log ring is too small for this run"
        log = split_log_by_tick(wb.meeting_log(sort=False))
        assert sorted(log) == sorted(meetings), "ticks with meetings differ"
        for t, m in meetings.items():
            assert np.array_equal(log[t], m), f"tick {t}: meetings differ ({len(log[t])} GPU vs {len(m)} oracle)"
        st = wb.stats()
        assert st["num_active"] == per_tick[-1]["num_active"]
        assert st["ref_drop_leaves"] == per_tick[-1]["ref_drop_leaves"]
        k, s = wb.active_pairs()
        ok, os_ = o.active()
        if not (cfg.record_contacts or cfg.geography) and w.min_meet_time != 0:
            keep = os_ != ticks - 1
            ok, os_ = ok[keep], os_[keep]
        assert np.array_equal(k, ok) and np.array_equal(s, os_), "active pairs after the last tick differ"
        b, lv, _ = wb.leaves()
        ob, olv, _ = o.leaves()
        assert np.array_equal(b, ob) and np.array_equal(lv, olv), "leaf boxes after the last tick differ"
        if extra:
            extra(wb, o)
    o.close()
    if own:
        trace.free()
    return {"per_tick": per_tick, "meetings": sum(len(m) for m in meetings.values())}


def run_positions_parity(xys: list, box, *, grid_capacity: int, reach: float = 2.0, min_meet: int = 5,
                         **cfg_kw) -> list[dict]:
    """The HIP path tick by tick on explicit positions (one FP64 AoS array
    per tick) against the oracle: every tick's meetings, active-pair and
    tested-pair counts, and the active pairs after the last tick.  Returns
    the GPU's per-tick statistics."""
    n = len(xys[0]) // 2
    o = Oracle(n, box, reach, grid_capacity, min_meet)
    kw = dict(reach_distance=reach, grid_capacity=grid_capacity, min_meet_time=min_meet)
    kw.update(cfg_kw)
    cfg = runtime.make_config(n, box, **kw)
    out = []
    with runtime.Workbench(cfg) as wb:
        for t, xy in enumerate(xys):
            o.tick(xy, t)
            m = wb.tick(xy, t)
            om = o.meetings()
            assert np.array_equal(m, om), f"tick {t}: meetings differ ({len(m)} GPU vs {len(om)} oracle)"
            st, ost = wb.stats(), o.stats()
            assert st["num_active"] == ost["num_active"], f"tick {t}: active pair counts differ"
            assert st["num_tests"] == ost["num_tests"], f"tick {t}: tested pairs differ"
            out.append(st)
        k, s = wb.active_pairs()
        ok, os_ = o.active()
        if not (cfg.record_contacts or cfg.geography) and min_meet != 0:
            keep = os_ != len(xys) - 1
            ok, os_ = ok[keep], os_[keep]
        assert np.array_equal(k, ok) and np.array_equal(s, os_), "active pairs after the last tick differ"
    o.close()
    return out
