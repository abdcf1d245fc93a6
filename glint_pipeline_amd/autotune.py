"""Autotuner: picks the launch and layout parameters of the contact-tracing
path per workload size class by timing short device-resident runs.

Knobs (none changes the emitted meetings except grid_capacity, which the
reference's candidate rule makes part of the result):
  join_tile      points per LDS tile of the distance join (256, 512)
  table_threads  workgroup size of the table update (256, 512, 1024)
  load_factor    table slots = next_pow2(distinct pairs per tick / load_factor); only the
                 stored pairs (those within reach at the tick before, about an eighth on
                 dense workloads) occupy slots, so values above 1 are valid
  precision      FP64 test, or FP32 guard band with FP64 recheck (both exact)
  grid_capacity  leaf capacity (64..1024); tuned only when the caller allows
                 it, and results are then compared with the oracle run at the
                 chosen capacity
The search is coordinate descent from the defaults: one knob at a time, the
others held at their best values so far.  Results per size class live in
tuned.json next to this module.
"""
from __future__ import annotations

import json
import math
import time
from dataclasses import dataclass, field
from pathlib import Path

from . import gen, runtime

TUNED_PATH = Path(__file__).resolve().parent / "tuned.json"

DEFAULTS = {"join_tile": 256, "table_threads": 512, "load_factor": 0.75, "precision": "fp32guard"}
SPACE = {"join_tile": [256, 512], "table_threads": [256, 512, 1024], "load_factor": [0.375, 0.75, 1.5, 3.0, 6.0],
         "precision": ["fp64", "fp32guard"]}
LEAF_CAPACITIES = [64, 128, 192, 256, 320, 384, 512, 1024]


def size_class(n: int) -> str:
    """Size classes of the batched cities: small < 200k <= medium < 800k <= large."""
    return "small" if n < 200_000 else ("medium" if n < 800_000 else "large")


def pairs_estimate(workload: str, n: int) -> float:
    """Distinct contact pairs per tick expected of a workload of n objects:
    they grow with the square of the hotspot density, about 65 per object at
    10M skewed (measured) and 0.06 per object for uniform walkers."""
    return {"skewed": 70.0 * n / 1e7, "plain_gpu": 0.1, "plain_cpu": 0.1, "batched": 0.1}[workload] * n


def table_slots(distinct: float, load_factor: float) -> int:
    return 1 << max(11, math.ceil(math.log2(max(distinct, 1.0) / load_factor)))


SCHEMAS = {"static": runtime.SCHEMA_STATIC, "dynamic": runtime.SCHEMA_DYNAMIC, "grid": runtime.SCHEMA_GRID_REBUILD}


def box_for(w: gen.Workload, schema: str = "dynamic") -> tuple:
    """The root box: the workload's, or one metre wider on every side for the
    grid rebuild, whose 16384^2 cell grid has no cell on the high y edge that
    the walkers reach."""
    lx, ly, hx, hy = w.box()
    return (lx - 1.0, ly - 1.0, hx + 1.0, hy + 1.0) if schema == "grid" else (lx, ly, hx, hy)


def config_for(w: gen.Workload, params: dict, distinct: float, graph_ticks: int = 8, **extra) -> runtime.Config:
    prec = runtime.PRECISION_FP32_GUARD if params.get("precision", "fp32guard") == "fp32guard" \
        else runtime.PRECISION_FP64
    n = w.num_objects
    schema = params.get("schema", "dynamic")
    return runtime.make_config(
        n, box_for(w, schema), reach_distance=w.reach_distance,
        grid_capacity=params.get("grid_capacity", w.grid_capacity), dynamic_schema=SCHEMAS[schema],
        min_meet_time=w.min_meet_time, precision=prec,
        num_meeting_buckets=table_slots(distinct, params.get("load_factor", 0.75)),
        unit_capacity=8 * n + 4096, meeting_capacity=max(1 << 20, 6 * n),
        contact_capacity=int(1.5 * distinct) + (1 << 20), graph_ticks=graph_ticks,
        join_tile=params.get("join_tile", 256), table_threads=params.get("table_threads", 0), **extra)


@dataclass
class Trial:
    params: dict
    ms_per_tick: float
    distinct: float


@dataclass
class TuneResult:
    best: dict
    best_ms: float
    trials: list = field(default_factory=list)


def measure(w: gen.Workload, params: dict, trace: runtime.DeviceBuffer, warmup: int, ticks: int,
            distinct: float) -> Trial:
    """Ticks warmup..warmup+ticks of a trace resident in device memory
    (tick k at 16*N*k bytes), timed between synchronisations."""
    n = w.num_objects
    with runtime.Workbench(config_for(w, params, distinct)) as wb:
        wb.run_ticks(trace.address(0), 0, warmup, latencies=False)
        wb.synchronize()
        t0 = time.perf_counter()
        wb.run_ticks(trace.address(16 * n * warmup), warmup, ticks, latencies=False)
        wb.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / ticks
        tot = wb.totals()
    return Trial(dict(params), ms, tot["max_distinct_tick"])


def probe_distinct(w: gen.Workload, trace: runtime.DeviceBuffer, ticks: int, hint: float) -> float:
    """Distinct pairs per tick, from a run sized for twice the hint."""
    with runtime.Workbench(config_for(w, DEFAULTS, 2.0 * hint, graph_ticks=0)) as wb:
        wb.run_ticks(trace.address(0), 0, ticks, latencies=False)
        return float(wb.totals()["max_distinct_tick"])


def device_trace(w: gen.Workload, ticks: int) -> runtime.DeviceBuffer:
    g = gen.Generator(w)
    buf = runtime.DeviceBuffer(16 * w.num_objects * ticks)
    buf.upload(g.positions(), 0)
    if ticks > 1:
        runtime.DeviceWalkers(g).walk(ticks - 1, buf, 16 * w.num_objects)
    return buf


def tune(w: gen.Workload, tune_capacity: bool = False, warmup: int = 3, ticks: int = 5,
         trace: runtime.DeviceBuffer | None = None, distinct_hint: float | None = None,
         fixed: dict | None = None, start: dict | None = None, log=print) -> TuneResult:
    """distinct_hint: expected distinct pairs per tick (sizes the probe run;
    default 4 per object); fixed: knobs the workload prescribes (not searched);
    start: the point the search starts from (default: DEFAULTS), e.g. the
    stored parameters, so that each knob is tried beside the others' best."""
    own = trace is None
    if own:
        trace = device_trace(w, warmup + ticks)
    try:
        hint = distinct_hint if distinct_hint else 4.0 * w.num_objects
        distinct = probe_distinct(w, trace, warmup + ticks, hint) * 1.25
        best = {**DEFAULTS, **(start or {}), **(fixed or {})}
        knobs = {k: v for k, v in SPACE.items() if k not in (fixed or {})}
        if tune_capacity:
            best["grid_capacity"] = (start or {}).get("grid_capacity", w.grid_capacity)
            knobs = {"grid_capacity": LEAF_CAPACITIES, **knobs}
        res = TuneResult(dict(best), float("inf"))
        seen = {}

        def run(p):
            key = json.dumps(p, sort_keys=True)
            if key not in seen:
                tr = measure(w, p, trace, warmup, ticks, distinct)
                seen[key] = tr
                res.trials.append({**tr.params, "ms_per_tick": round(tr.ms_per_tick, 3)})
                log(f"[autotune] {w.name}: {tr.params} {tr.ms_per_tick:.3f} ms/tick")
            return seen[key].ms_per_tick

        res.best_ms = run(best)
        for k, values in knobs.items():
            for v in values:
                cand = {**best, k: v}
                ms = run(cand)
                if ms < res.best_ms:
                    best, res.best_ms = cand, ms
        res.best = best
        return res
    finally:
        if own:
            trace.free()


def load_table() -> dict:
    if TUNED_PATH.exists():
        return json.loads(TUNED_PATH.read_text())
    return {}


def save_table(table: dict) -> None:
    TUNED_PATH.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")


def params_for(key: str) -> dict:
    """Tuned parameters stored under key (a size class or a workload name),
    or the defaults."""
    entry = load_table().get(key)
    return dict(entry["params"]) if entry else dict(DEFAULTS)
