"""Shared runs of the sorted key-value output: a small seeded city traced by
the oracle (and optionally by the device path) with flushes at given ticks.

Coordinates stay below 820 m so that float_to_uint of every box stays under
the reference's object-box initial low (1e8), where the object boxes are
meaningful (cuda_util.cuh:115-119, geometry.cu:1254-1263)."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from glint_pipeline_amd import gen
from oracle.oracle import Oracle


@dataclass
class KvCase:
    n: int = 2500
    side: float = 500.0
    ticks: int = 60
    seed: int = 5
    grid_capacity: int = 16
    min_meet: int = 3
    split: int = 4
    max_meet: int = 4096
    area_limit: float = 1e9
    flush_at: tuple = (35, 59)
    hotspots: int = 0

    def workload(self) -> gen.Workload:
        w = gen.uniform_city(self.n, self.ticks, self.side, seed=self.seed, grid_capacity=self.grid_capacity,
                             min_meet_time=self.min_meet)
        if self.hotspots:
            w.num_hotspots = self.hotspots
            w.hotspot_fraction = 0.7
            w.hotspot_sigma = 40.0
        return w


@dataclass
class KvRun:
    trace: np.ndarray                     # (ticks, n, 2)
    meetings: list = field(default_factory=list)   # per tick: (m, 4) u32
    flushes: list = field(default_factory=list)    # per flush: dict from kv_flush


def run_oracle(case: KvCase) -> KvRun:
    w = case.workload()
    g = gen.Generator(w)
    o = Oracle(w.num_objects, w.box(), w.reach_distance, w.grid_capacity, w.min_meet_time)
    o.kv_enable(split_num=case.split, max_meet_time=case.max_meet, area_limit=case.area_limit)
    trace = np.zeros((case.ticks, case.n, 2))
    run = KvRun(trace)
    for t in range(case.ticks):
        xy = (g.positions() if t == 0 else g.step()).copy()
        trace[t] = xy.reshape(-1, 2)
        o.tick(xy, t)
        run.meetings.append(o.meetings())
        if t in case.flush_at:
            run.flushes.append(o.kv_flush())
    o.close()
    return run


def run_device(case: KvCase, kv_capacity: int = 1 << 16, graph_ticks: int = 0) -> KvRun:
    from glint_pipeline_amd import runtime
    w = case.workload()
    g = gen.Generator(w)
    cfg = runtime.make_config(w.num_objects, w.box(), reach_distance=w.reach_distance,
                              grid_capacity=w.grid_capacity, min_meet_time=w.min_meet_time,
                              max_meet_time=case.max_meet, kv_capacity=kv_capacity, kv_split_num=case.split,
                              kv_area_limit=case.area_limit, graph_ticks=graph_ticks)
    trace = np.zeros((case.ticks, case.n, 2))
    run = KvRun(trace)
    with runtime.Workbench(cfg) as wb:
        if graph_ticks:
            d = runtime.DeviceBuffer(case.ticks * case.n * 16)
            for t in range(case.ticks):
                trace[t] = (g.positions() if t == 0 else g.step()).reshape(-1, 2)
            d.upload(trace)
            t0 = 0
            for t1 in list(case.flush_at):
                wb.run_ticks(d.address(16 * case.n * t0), t0, t1 + 1 - t0)
                run.flushes.append(wb.kv_flush())
                t0 = t1 + 1
            d.free()
        else:
            for t in range(case.ticks):
                xy = (g.positions() if t == 0 else g.step()).copy()
                trace[t] = xy.reshape(-1, 2)
                run.meetings.append(wb.tick(xy, t))
                if t in case.flush_at:
                    run.flushes.append(wb.kv_flush())
    return run


def fields(keys: np.ndarray):
    """(sid, oid, target, duration, end offset) of (lo, hi) key pairs."""
    lo = keys[:, 0].astype(np.uint64)
    hi = keys[:, 1].astype(np.uint64)
    sid = (hi >> np.uint64(52)).astype(np.int64)
    oid = ((hi >> np.uint64(26)) & np.uint64((1 << 26) - 1)).astype(np.int64)
    tgt = (hi & np.uint64((1 << 26) - 1)).astype(np.int64)
    dur = (lo >> np.uint64(32)).astype(np.int64)
    end = (lo & np.uint64(0xFFFFFFFF)).astype(np.int64)
    return sid, oid, tgt, dur, end


def f2u(v):
    """float_to_uint in float32 arithmetic."""
    v = np.asarray(v, dtype=np.float32)
    return ((v + np.float32(180)) * np.float32(100000)).astype(np.uint32)


def u2f(u):
    return np.asarray(u, dtype=np.uint32).astype(np.float32) / np.float32(100000) - np.float32(180)
