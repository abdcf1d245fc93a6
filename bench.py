"""Benchmark of the contact-tracing hot path on one MI355X.

A step is one tick of the path (partition -> refine -> evaluate -> schema
update) over every object.  Default workload: the skewed-density scenario,
10M objects in a 30 km box, 80% starting in 64 Gaussian hotspots (sigma 200 m,
Zipf s=1.1 weights), 20% uniform, all walking a 50 m Manhattan grid at
U[1,15) m/s; reach 2 m, minimum meeting 5 ticks; FP32 guard-band test with
FP64 recheck, 8 ticks per captured hipGraph, leaf capacity, join tile and
table load factor from the autotuner (tuned.json, or --autotune).  Positions
for every tick are generated on the device before the timed region starts.
--workload batched times 32 independent cities of uneven size in one batch.

Prints one JSON line (rank 0) with object-updates/s, p50/p99 tick latency,
the roofline figure of the dominant kernel (HIP events around every launch
on the context's stream) and the CPU oracle timed on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "object-updates/sec and p99 tick latency at 10M objects, 1 MI355X, exact contacts"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_for(name: str, objects: int | None):
    from glint_pipeline_amd import gen
    if name == "skewed":
        w = gen.skewed_density(objects or 10_000_000, 300)
        desc = ("skewed density: {n} objects, 30 km box, 80% in 64 Gaussian hotspots (sigma 200 m, Zipf s=1.1), "
                "20% uniform, Manhattan walkers U[1,15) m/s, r=2 m, min meeting 5, leaf capacity 200")
    elif name == "plain_gpu":
        w = gen.plain_gpu_form()
        if objects:
            w.num_objects = objects
        desc = "plain GPU form: {n} objects, 30 km box, uniform Manhattan walkers, r=2 m, min 5, leaf capacity 200"
    elif name == "plain_cpu":
        w = gen.plain_cpu_reference()
        if objects:
            w.num_objects = objects
        desc = "plain CPU reference: {n} objects, 10 km box, uniform Manhattan walkers, r=2 m, min 5, capacity 200"
    else:
        raise SystemExit(f"unknown workload {name}")
    return w, desc.format(n=w.num_objects)


def algorithmic_bytes(kernel: str, n: int, per_tick: dict, precision: int) -> float:
    """Algorithmic HBM bytes of one launch of `kernel` (one tick), from the
    per-tick figures: N objects, S contacts kept in the meeting table (those
    within reach at the tick before or found through a border unit), M
    meetings, U border units."""
    S, M, U = per_tick["stored"], per_tick["meetings"], per_tick["units"]
    coord = 8 if precision else 16          # FP32 offsets or FP64 coordinates per object
    if kernel == "k_join":
        # every sorted object once (coordinates, previous FP64 position, id),
        # each border unit (id + xy), one 8 B key per kept contact out
        return (coord + 16.0 + 4.0) * n + 24.0 * U + 8.0 * S
    if kernel == "k_table":
        # the previous generation's entries (8 B key + 4 B start; as many as
        # the tick before kept, taken equal to S), every kept contact's key
        # read and its start written, meetings out
        return 12.0 * S + 12.0 * S + 16.0 * M
    if kernel == "k_bucket_fine":
        return 24.0 * S                       # segment read twice (histogram, tiles), written once
    if kernel == "k_bucket_scatter":
        return 16.0 * S
    if kernel == "k_bucket_count":
        return 8.0 * S
    if kernel == "k_assign":
        return 16.0 * n + 8.0 * n             # xy in, leaf and rank out
    if kernel == "k_scatter":
        return 16.0 * n + 8.0 * n + 20.0 * n  # xy, leaf/rank in; sorted xy + id out
    if kernel == "k_lookup":
        return 8.0 * per_tick["border_objects"] + 16.0 * per_tick["border_objects"] + 8.0 * U
    return 0.0


def profiled_traffic(kernel: str, workload: str, precision: int, objects: int) -> tuple[float | None, str]:
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE runs scaled by a
    plain copy's counters), and where the figure comes from.  A summary made
    from other kernel sources than the ones built now is not used."""
    from glint_pipeline_amd.build import source_digest
    tag = f"traffic_{workload}_{objects // 1_000_000}m_{'fp32guard' if precision else 'fp64'}.txt"
    f = ROOT / "profiles" / tag
    if not f.exists():
        return None, f"no counter summary profiles/{tag}"
    found, sha = None, None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "TRAFFIC_BYTES_PER_LAUNCH" and parts[1] == kernel:
            found = float(parts[2])
        if len(parts) == 2 and parts[0] == "SOURCE_SHA":
            sha = parts[1]
    cur = source_digest()
    if sha != cur:
        return None, f"profiles/{tag} measured kernel sources {sha}, built sources are {cur}"
    return found, (f"profiles/{tag} (FETCH_SIZE + WRITE_SIZE per launch averaged over the ticks the summary "
                   f"names, copy-calibrated; sources {cur})")


VALU_PEAK = 256 * 2.4e9   # wave64 VALU instructions per second: one per CU and cycle, 256 CUs at 2.4 GHz


def profiled_valu(kernel: str, workload: str, precision: int,
                  objects: int) -> tuple[float | None, float | None, str]:
    """VALU wave instructions per launch of `kernel` from the committed
    rocprofv3 counter summary (SQ_INSTS_VALU), if it measured the kernel
    sources built now, with the pair slots per launch of the ticks it
    measured (None if not recorded)."""
    from glint_pipeline_amd.build import source_digest
    tag = f"valu_{workload}_{objects // 1_000_000}m_{'fp32guard' if precision else 'fp64'}.txt"
    f = ROOT / "profiles" / tag
    if not f.exists():
        return None, None, f"no counter summary profiles/{tag}"
    found, slots, sha = None, None, None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "VALU_INSTS_PER_LAUNCH" and parts[1] == kernel:
            found = float(parts[2])
        if len(parts) == 3 and parts[0] == "PAIR_SLOTS_PER_LAUNCH" and parts[1] == kernel:
            slots = float(parts[2])
        if len(parts) == 2 and parts[0] == "SOURCE_SHA":
            sha = parts[1]
    cur = source_digest()
    if sha != cur:
        return None, None, f"profiles/{tag} measured kernel sources {sha}, built sources are {cur}"
    return found, slots, f"profiles/{tag} (SQ_INSTS_VALU; sources {cur})"


# the tick's kernels by the reference's stages (filter = partition and
# lookup, reachability = join and the grouping of contacts by table region,
# update_meetings = the table update, update_schema); the scans serve all three
STAGE_OF_KERNEL = {
    "k_fill_u32": "filter", "k_reset_tick": "filter", "k_regions": "filter", "k_assign": "filter",
    "k_scan_objects": "filter", "k_scatter": "filter", "k_fill_ucount": "filter", "k_lookup": "filter",
    "k_scan_units": "filter", "k_unit_scatter": "filter", "k_grid_cells": "filter", "k_grid_pyramid": "filter",
    "k_grid_top": "filter", "k_grid_fine_zero": "filter", "k_grid_fine": "filter", "k_grid_level": "filter",
    "k_join": "reachability", "k_bucket_count": "reachability", "k_scan_write": "reachability",
    "k_bucket_scatter": "reachability", "k_bucket_fine": "reachability",
    "k_table": "update_meetings", "k_table_reshaped": "update_meetings", "k_pos32": "update_meetings",
    "k_log_append": "update_meetings",
    "k_collect": "update_schema", "k_split": "update_schema", "k_merge": "update_schema",
    "k_schema_done": "update_schema",
}


def gpu_stages(ktimes: dict, ticks: int) -> dict:
    out = {"filter": 0.0, "reachability": 0.0, "update_meetings": 0.0, "update_schema": 0.0, "other": 0.0}
    for name, (ms, _cnt) in ktimes.items():
        out[STAGE_OF_KERNEL.get(name, "other")] += ms / ticks
    return {k: round(v, 3) for k, v in out.items()}


def cpu_baseline(w, ticks: int, threads: int, min_seconds: float = 10.0, max_ticks: int = 5000,
                 schema: str = "dynamic", geo: "LonLat | None" = None) -> dict:
    """The CPU oracle (C + OpenMP) on the first ticks of the same workload:
    at least `ticks` ticks, and more (the walk continues) until about
    `min_seconds` of CPU work is timed.  With geo, the positions are mapped
    to degrees and the oracle runs the geography predicates."""
    from glint_pipeline_amd import autotune, gen
    from oracle.oracle import Oracle
    g = gen.Generator(w)
    box = geo.box(w) if geo else autotune.box_for(w, schema)
    o = Oracle(w.num_objects, box, w.reach_distance, w.grid_capacity, w.min_meet_time,
               num_threads=threads, dynamic_schema=autotune.SCHEMAS[schema], geography=geo is not None)
    conv = geo.to_degrees if geo else (lambda xy: xy)
    t_build0 = time.perf_counter()
    o.tick(conv(g.positions()), 0)    # tick 0 builds the tree (not timed, as on the GPU side)
    t_build = time.perf_counter() - t_build0
    tt = 0.0
    t = 0
    stages = {"filter": 0.0, "reachability": 0.0, "update_meetings": 0.0, "update_schema": 0.0}
    while t < max_ticks and (t < ticks or tt < min_seconds):
        t += 1
        xy = conv(g.step())
        t0 = time.perf_counter()
        o.tick(xy, t)
        tt += time.perf_counter() - t0
        st = o.stats()   # the oracle's own timers of its stages (the reference's profiler, workbench.h:19-53)
        for k, f in (("filter", "t_filter"), ("reachability", "t_refine"), ("update_meetings", "t_evaluate"),
                     ("update_schema", "t_schema")):
            stages[k] += st[f]
    st = o.stats()
    o.close()
    return {"seconds": tt, "ticks": t, "tick0_seconds": t_build, "contacts_last_tick": st["num_contacts"],
            "stages_ms_per_tick": {k: round(v * 1e3 / t, 3) for k, v in stages.items()}}


class LonLat:
    """A walk in metres mapped to longitude/latitude degrees near (lon0,
    lat0), the input of the reference's geography=true predicates; the reach
    and the split rules stay in metres.  The degrees per metre are the
    reference's expressions (util.h:57-66, PI = 3.14159265)."""

    def __init__(self, lon0: float = -87.63, lat0: float = 41.88):
        self.lon0, self.lat0 = lon0, lat0
        self.fx = 360.0 / (math.sin((90.0 - abs(lat0)) * 3.14159265 / 180.0) * 40076.0 * 1000.0)
        self.fy = 360.0 / (40076.0 * 1000.0)

    def to_degrees(self, xy: np.ndarray) -> np.ndarray:
        ll = np.array(xy, dtype=np.float64).reshape(-1, 2)
        ll[:, 0] = self.lon0 + ll[:, 0] * self.fx
        ll[:, 1] = self.lat0 + ll[:, 1] * self.fy
        return ll.reshape(np.shape(xy))

    def box(self, w) -> tuple:
        # the walk's box with 1 m of margin, in degrees
        return (self.lon0 - self.fx, self.lat0 - self.fy, self.lon0 + (w.side + 1) * self.fx,
                self.lat0 + (w.side + 1) * self.fy)


WORKLOAD_SPEC = {
    # config 3: FP32 guard band + FP64 recheck, 8 ticks per graph, tuned leaf capacity
    "skewed": {"precision": "fp32guard", "graph_ticks": 8, "tuned": "skewed_10m"},
    # config 2: FP64 throughout, one call sequence per tick
    "plain_gpu": {"precision": "fp64", "graph_ticks": 0, "tuned": None},
    "plain_cpu": {"precision": "fp64", "graph_ticks": 0, "tuned": None},
}


def distinct_estimate(workload: str, n: int) -> float:
    from glint_pipeline_amd import autotune
    return autotune.pairs_estimate(workload, n)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=96)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="skewed", choices=["skewed", "plain_gpu", "plain_cpu", "batched", "serving"])
    ap.add_argument("--history", type=int, default=3600, help="serving: ticks of history before the timed ticks")
    ap.add_argument("--no-pace", action="store_true",
                    help="serving: run ticks back to back with 50 queries per tick (a stress test) instead of "
                         "one tick per wall-clock second with Poisson arrivals at 50/s")
    ap.add_argument("--objects", type=int, default=None)
    ap.add_argument("--snapshot", action="store_true",
                    help="after the timed run, time a snapshot of the tick state to host memory and its restore")
    ap.add_argument("--geography", action="store_true",
                    help="positions mapped to lon/lat degrees; the reference's geography=true predicates (FP64)")
    ap.add_argument("--precision", default=None, choices=["fp64", "fp32guard"],
                    help="default: the workload's configuration (skewed: fp32guard; plain forms: fp64)")
    ap.add_argument("--graph-ticks", type=int, default=None, help="ticks per captured hipGraph (default: per workload)")
    ap.add_argument("--grid-capacity", type=int, default=None, help="leaf capacity (default: tuned or 200)")
    ap.add_argument("--join-tile", type=int, default=None, choices=[256, 512])
    ap.add_argument("--table-threads", type=int, default=None, choices=[256, 512, 1024],
                    help="workgroup size of the table update (default: tuned or 256)")
    ap.add_argument("--load-factor", type=float, default=None)
    ap.add_argument("--schema", default=None, choices=["dynamic", "grid", "static"],
                    help="tree maintenance: split/merge (default), rebuilt from the 16384^2 grid every tick, "
                         "or the first tick's tree kept")
    ap.add_argument("--autotune", action="store_true", help="tune leaf capacity, join tile, load factor and "
                                                             "precision on this workload before timing")
    ap.add_argument("--save-tuned", action="store_true", help="store the autotuned parameters in tuned.json")
    ap.add_argument("--no-profile", action="store_true",
                    help="run the second (verification) run without HIP events around every kernel")
    ap.add_argument("--cpu-ticks", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip timing the plain per-tick FP64 form (host positions, one call sequence per tick) "
                         "against the batched guard-band form on the same inputs")
    ap.add_argument("--kv", action="store_true",
                    help="also time the ticks with the sorted key-value output on (pair boxes, meeting log) "
                         "and its flushes at the reference's kv_restriction")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        return 2

    from glint_pipeline_amd import autotune, gen, runtime

    if runtime.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "serving":
        return bench_serving(args)
    w, desc = workload_for(args.workload, args.objects)
    N = w.num_objects
    W, K = args.warmup, args.steps
    spec = WORKLOAD_SPEC[args.workload]
    params = autotune.params_for(spec["tuned"]) if spec["tuned"] else dict(autotune.DEFAULTS)
    params["precision"] = spec["precision"]     # prescribed by the configuration, not tuned
    if spec["tuned"] and "grid_capacity" not in params:
        params["grid_capacity"] = w.grid_capacity
    graph_ticks = spec["graph_ticks"] if args.graph_ticks is None else args.graph_ticks
    distinct = distinct_estimate(args.workload, N)

    # ---- inputs: host init of tick 0, device walk for the rest, all in HBM
    t0 = time.perf_counter()
    g = gen.Generator(w)
    dw = runtime.DeviceWalkers(g)
    total_ticks = W + K
    trace = runtime.DeviceBuffer(16 * N * total_ticks)
    trace.upload(g.positions(), 0)
    if total_ticks > 1:
        dw.walk(total_ticks - 1, trace, 16 * N)
    log(f"[bench] generated {total_ticks} ticks of positions on the device in {time.perf_counter() - t0:.1f} s")
    geo = LonLat() if args.geography else None
    if geo:
        # every tick's positions mapped to degrees in place (FP64 predicates:
        # the guard band is defined on planar offsets)
        host = np.empty((N, 2), dtype=np.float64)
        for t in range(total_ticks):
            trace.download(host, 16 * N * t)
            trace.upload(geo.to_degrees(host), 16 * N * t)
        desc += f"; positions in lon/lat degrees near ({geo.lon0}, {geo.lat0}), the geography predicates"
        args.no_forms = True

    tuning = None
    if args.autotune:
        t0 = time.perf_counter()
        res = autotune.tune(w, tune_capacity=bool(spec["tuned"]), warmup=min(3, W), ticks=4, trace=trace,
                            distinct_hint=distinct, fixed={"precision": spec["precision"]}, start=params, log=log)
        params.update(res.best)
        tuning = {"best": res.best, "best_ms_per_tick": round(res.best_ms, 3), "trials": res.trials,
                  "seconds": round(time.perf_counter() - t0, 1)}
        if args.save_tuned and spec["tuned"]:
            table = autotune.load_table()
            table[spec["tuned"]] = {"params": res.best, "ms_per_tick": round(res.best_ms, 3),
                                    "workload": desc, "trials": res.trials}
            autotune.save_table(table)
    for k, v in (("precision", args.precision), ("grid_capacity", args.grid_capacity), ("join_tile", args.join_tile),
                 ("load_factor", args.load_factor), ("table_threads", args.table_threads), ("schema", args.schema)):
        if v is not None:
            params[k] = v
    if geo:
        params["precision"] = "fp64"
    precision = runtime.PRECISION_FP32_GUARD if params["precision"] == "fp32guard" else runtime.PRECISION_FP64
    w.grid_capacity = params.get("grid_capacity", w.grid_capacity)
    desc = desc.replace("leaf capacity 200", f"leaf capacity {w.grid_capacity}")

    def make_cfg(prec: int, graph: int, **extra):
        p = dict(params)
        p["precision"] = "fp32guard" if prec == runtime.PRECISION_FP32_GUARD else "fp64"
        c = autotune.config_for(w, p, distinct, graph_ticks=graph, geography=geo is not None, **extra)
        if geo:
            c.low_x, c.low_y, c.high_x, c.high_y = geo.box(w)
        return c

    cfg = make_cfg(precision, graph_ticks)
    slots = int(cfg.num_meeting_buckets)
    log(f"[bench] workload: {desc}; params {params}; graph ticks {graph_ticks}; warmup {W}, timed {K}; "
        f"table slots 2^{int(math.log2(slots))}")
    wb = runtime.Workbench(cfg)
    t0 = time.perf_counter()
    if W:
        wb.run_ticks(trace.address(0), 0, W, latencies=False)
    wb.synchronize()
    log(f"[bench] warmup ({W} ticks incl. tree build) {time.perf_counter() - t0:.2f} s; stats {wb.stats()}")
    wb.totals(reset=True)

    # ---- timed region
    wb.synchronize()
    t0 = time.perf_counter()
    tick_ms = wb.run_ticks(trace.address(16 * N * W), W, K, latencies=True)
    wb.synchronize()
    elapsed = time.perf_counter() - t0
    totals = wb.totals()
    stats = wb.stats()
    log(f"[bench] timed {K} ticks in {elapsed:.3f} s; totals {totals}")
    snap = None
    if args.snapshot:
        # the tick-boundary state to host memory and into a fresh context of
        # the same configuration (the serving loop's dump and reload)
        t0 = time.perf_counter()
        buf = wb.snapshot()
        t_snap = time.perf_counter() - t0
        wr = runtime.Workbench(make_cfg(precision, graph_ticks))
        t0 = time.perf_counter()
        wr.restore(buf)
        wr.synchronize()
        t_rest = time.perf_counter() - t0
        same = wr.active_pairs()[0].size == wb.active_pairs()[0].size
        wr.close()
        snap = {"bytes": int(buf.size), "snapshot_ms": round(t_snap * 1e3, 1), "restore_ms": round(t_rest * 1e3, 1),
                "host_GBps": round(buf.size / max(t_snap, 1e-9) / 1e9, 2), "same_active_pairs": bool(same)}
        log(f"[bench] snapshot {snap}")
        del buf
    wb.close()

    # ---- verification and per-kernel times: a second run of the same inputs
    # without graph capture (events cannot sit inside a replayed graph), HIP
    # events around every launch on the context's stream during ticks
    # W..W+K-1, the meeting log on.  Its totals over the timed ticks must equal
    # the graph run's, and where the committed oracle fixture covers this
    # workload, every fixture tick's meetings must equal the oracle's.
    fixture = None if geo else headline_fixture(args.workload, w, params)
    fx_ticks = min(len(fixture["ticks"]), W + K) if fixture else 0
    wv = runtime.Workbench(make_cfg(precision, 0, meeting_log_capacity=(64 << 20) if fixture else 0))
    verify = {"graph_vs_plain": None, "fixture_ticks": 0}
    mismatches = []
    done = 0

    def advance(upto: int) -> None:
        nonlocal done
        if upto <= done:
            return
        wv.run_ticks(trace.address(16 * N * done), done, upto - done, latencies=False)
        done = upto

    def start_timed() -> None:
        wv.synchronize()
        wv.totals(reset=True)
        if not args.no_profile:
            wv.set_kernel_profiling(True)
            wv.kernel_times(reset=True)

    ktimes = {}
    if W == 0:
        start_timed()
    for cut in sorted({W, fx_ticks, W + K} - {0}):
        if done < W <= cut:
            advance(W)
            start_timed()
        advance(cut)
        if fixture and cut == fx_ticks:
            wv.synchronize()
            mismatches += check_fixture(wv, fixture, fx_ticks)
            verify["fixture_ticks"] = fx_ticks
    wv.synchronize()
    gpu_stage_ms = None
    if not args.no_profile:
        ktimes = wv.kernel_times()
        gpu_stage_ms = gpu_stages(ktimes, K)
        wv.set_kernel_profiling(False)
        ptick = sum(ms for ms, cnt in ktimes.values()) / K
        for name, (ms, cnt) in sorted(ktimes.items(), key=lambda kv: -kv[1][0]):
            if cnt:
                log(f"[bench]   {name:16s} {ms / K:9.3f} ms/tick  {cnt // K:4d} launches/tick  {ms / cnt:9.3f} ms/launch")
        log(f"[bench]   kernels sum {ptick:.3f} ms/tick (events on every launch, no graph, ticks {W}..{W + K - 1})")
    vt = wv.totals()
    same = [k for k in ("ticks", "contacts", "distinct_pairs", "meetings", "stored", "units", "border_objects",
                        "rechecks", "ref_drop_objects") if vt[k] != totals[k]]
    vs = wv.stats()
    if vs["num_active"] != stats["num_active"] or vs["num_leaves"] != stats["num_leaves"]:
        same.append("last tick's active pairs or leaves")
    verify["graph_vs_plain"] = not same
    if same:
        mismatches.append(f"graph run and plain run differ in {same}")
    wv.close()
    log(f"[bench] verification {verify}; {len(mismatches)} mismatches")

    value = N * K / elapsed
    per_tick = {k: totals[k] / max(1, totals["ticks"]) for k in
                ("contacts", "contacts_border", "meetings", "units", "border_objects", "distinct_pairs", "stored",
                 "ref_drop_leaves", "ref_drop_objects")}
    roofline = None
    if ktimes:
        dom, (dms, dcnt) = max(((k, v) for k, v in ktimes.items() if v[1]), key=lambda kv: kv[1][0])
        abytes = algorithmic_bytes(dom, N, per_tick, precision)
        avg_s = dms / dcnt / 1e3
        achieved = abytes / avg_s / 1e9 if abytes and avg_s > 0 else 0.0
        traffic, tnote = profiled_traffic(dom, args.workload, precision, N)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tnote,
                    "kernel": dom, "kernel_ms_per_launch": round(dms / dcnt, 4),
                    "algorithmic_bytes_per_launch": abytes,
                    "kernel_share_of_tick": round(dms / K / (elapsed * 1e3 / K), 4),
                    "measured": f"HIP events on every launch over ticks {W}..{W + K - 1}, ungraphed run of the same "
                                f"inputs"}
        if dom == "k_join":
            # the join's own bound is instruction issue: VALU wave instructions
            # per launch (committed counter summary) against the issue peak,
            # and the candidate pairs it tests per second
            vi, vslots, vnote = profiled_valu(dom, args.workload, precision, N)
            slots = totals["evaluated"] / max(1, totals["ticks"])
            if vi and vslots:
                # the profile's ticks differ from this run's: the instruction
                # count scales with the pair slots evaluated
                vi = vi * slots / vslots
                vnote += f"; scaled by this run's pair slots per launch ({slots:.4g} against {vslots:.4g})"
            roofline["compute"] = {
                "bound": "valu issue", "unit": "wave instructions/s", "peak": VALU_PEAK,
                "insts_per_launch": round(vi) if vi else None, "achieved": round(vi / avg_s, 1) if vi else None,
                "frac": round(vi / avg_s / VALU_PEAK, 4) if vi else None, "source": vnote,
                "pair_tests_per_s": round(totals["tests"] / max(1, totals["ticks"]) / avg_s, 1),
                "pair_slots_evaluated_per_s": round(totals["evaluated"] / max(1, totals["ticks"]) / avg_s, 1)}

    # ---- the plain FP64 per-tick form on the same inputs: positions copied
    # from host memory every tick, the four stage calls, a host sync per tick
    forms = None
    if not args.no_forms:
        forms = {}
        host = np.zeros(2 * N, dtype=np.float64)
        wp = runtime.Workbench(make_cfg(runtime.PRECISION_FP64, 0))
        for t in range(W):
            trace.download(host, 16 * N * t)
            wp.tick(host, t)
        wp.synchronize()
        wp.totals(reset=True)
        lat = []
        for t in range(W, W + K):
            trace.download(host, 16 * N * t)       # the caller's positions, in host memory (not timed)
            t1 = time.perf_counter()
            wp.set_points(host, t)                 # host-to-device copy every tick, as the reference does
            wp.filter()
            wp.reachability()
            wp.update_meetings(copy_out=False)
            wp.update_schema()
            wp.synchronize()
            lat.append((time.perf_counter() - t1) * 1e3)
        pt = wp.totals()
        wp.close()
        if pt["meetings"] != totals["meetings"] or pt["distinct_pairs"] != totals["distinct_pairs"]:
            mismatches.append("plain FP64 form and guard-band form differ")
        forms["plain_fp64_per_tick_ms"] = round(float(np.sum(lat)) / K, 3)
        forms["plain_p99_ms"] = round(float(np.percentile(lat, 99)), 3)
        if precision == runtime.PRECISION_FP32_GUARD and graph_ticks == 8:
            forms["batched_guard_graph8_ms"] = round(elapsed * 1e3 / K, 3)   # the timed run is that form
            forms["batched_p99_ms"] = round(float(np.percentile(tick_ms, 99)), 3)
        else:
            wq = runtime.Workbench(make_cfg(runtime.PRECISION_FP32_GUARD, 8))
            wq.run_ticks(trace.address(0), 0, W, latencies=False)
            wq.synchronize()
            t0 = time.perf_counter()
            ms = wq.run_ticks(trace.address(16 * N * W), W, K, latencies=True)
            wq.synchronize()
            forms["batched_guard_graph8_ms"] = round((time.perf_counter() - t0) * 1e3 / K, 3)
            forms["batched_p99_ms"] = round(float(np.percentile(ms, 99)), 3)
            wq.close()
        forms["speedup"] = round(forms["plain_fp64_per_tick_ms"] / forms["batched_guard_graph8_ms"], 3)
        forms["same_meetings"] = pt["meetings"] == totals["meetings"]
        log(f"[bench] forms {forms}")

    kv = None
    if args.kv:
        restriction = 2 * 33554432                 # kv_restriction at G_bytes = 2 (config.h:53,80)
        wk = runtime.Workbench(make_cfg(precision, graph_ticks, kv_capacity=2 * restriction,
                                        kv_area_limit=1e30))   # keep every object in the files
        wk.run_ticks(trace.address(0), 0, W, latencies=False)
        wk.synchronize()
        flush_ms, keys = [], []
        tick_s = 0.0
        t = W
        while t < W + K:
            step = min(8, W + K - t)
            t1 = time.perf_counter()
            wk.run_ticks(trace.address(16 * N * t), t, step, latencies=False)
            wk.synchronize()
            tick_s += time.perf_counter() - t1
            t += step
            if wk.kv_pending() > restriction or t == W + K:
                t1 = time.perf_counter()
                fl = wk.kv_flush(copy=False)     # organised on the device; host copies not timed
                flush_ms.append((time.perf_counter() - t1) * 1e3)
                keys.append(fl["summary"]["kept_keys"] + fl["summary"]["oversize_keys"])
        wk.close()
        kv = {"tick_ms": round(tick_s * 1e3 / K, 3), "flushes": len(flush_ms),
              "flush_ms": [round(v, 1) for v in flush_ms], "keys_per_flush": keys,
              "per_tick_ms_with_flushes": round((tick_s * 1e3 + sum(flush_ms)) / K, 3)}
        log(f"[bench] key-value output {kv}")

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        t0 = time.perf_counter()
        cb = cpu_baseline(w, args.cpu_ticks, threads, schema=params.get("schema", "dynamic"), geo=geo)
        log(f"[bench] cpu baseline {cb} (wall {time.perf_counter() - t0:.1f} s)")
        cpu = {"value": round(N * cb["ticks"] / cb["seconds"], 1), "unit": "object-updates/s", "cores": threads,
               "kind": "port", "cpu": cpu_model(), "stages_ms_per_tick": cb["stages_ms_per_tick"],
               "sample": f"oracle (C, OpenMP) ticks 1..{cb['ticks']} of the same workload after its tick-0 tree "
                         f"build; {cb['seconds']:.2f} s timed"}

    lat = np.asarray(tick_ms, dtype=np.float64)
    worst = int(np.argmax(lat))
    p99_note = {"tick": W + worst, "ms": round(float(lat[worst]), 3), "index_in_run": worst,
                "index_in_graph": worst % graph_ticks if graph_ticks else None,
                "ticks_over_1.25x_p50": int(np.count_nonzero(lat > 1.25 * np.percentile(lat, 50)))}
    if mismatches:
        print(json.dumps({"metric": METRIC, "value": None, "unit": "object-updates/s", "n_gpus": 1,
                          "error": "results not verified: " + "; ".join(mismatches), "verification": verify}),
              flush=True)
        trace.free()
        return 3
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "object-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(elapsed * 1e3 / K, 3),
        "p50_tick_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_tick_ms": round(float(np.percentile(lat, 99)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == runtime.PRECISION_FP64 else "f32+f64",
        "data": "synthetic (seeded generator, positions resident in HBM before timing)",
        "config": {"workload": desc, "objects": N, "first_tick": W, "ticks": K, "graph_ticks": graph_ticks,
                   "params": params, "table_slots_last_tick": stats["table_slots"],
                   "contacts_per_tick": round(per_tick["contacts"]),
                   "stored_per_tick": round(per_tick["stored"]), "meetings_per_tick": round(per_tick["meetings"], 1),
                   "leaves": stats["num_leaves"], "max_leaf_count": totals["max_leaf_count"],
                   "ref_drop_leaves_per_tick": round(per_tick["ref_drop_leaves"], 1),
                   "ref_drop_objects_per_tick": round(per_tick["ref_drop_objects"], 1)},
        "worst_tick": p99_note,
        "stages_ms_per_tick": gpu_stage_ms,   # kernel time by the reference's stages (ungraphed run, HIP events)
        "verification": verify,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if tuning:
        out["autotune"] = tuning
    if forms:
        out["forms"] = forms
    if kv:
        out["kv_output"] = kv
    if snap:
        out["snapshot"] = snap
    print(json.dumps(out), flush=True)
    trace.free()
    return 0


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


FIXTURE = ROOT / "tests" / "golden" / "skewed_10m_fixture.json"


def headline_fixture(workload: str, w, params: dict) -> dict | None:
    """The committed oracle fixture of this workload, if it is the one the
    fixture was written for (10M skewed objects at its leaf capacity)."""
    if workload != "skewed" or not FIXTURE.exists():
        return None
    fx = json.loads(FIXTURE.read_text())
    c = fx["case"]
    if (w.num_objects, w.grid_capacity, w.reach_distance, w.min_meet_time) != \
            (c["objects"], c["grid_capacity"], c["reach"], c["min_meet"]) or params.get("schema", "dynamic") != "dynamic":
        return None
    return fx


def check_fixture(wb, fx: dict, ticks: int) -> list[str]:
    """Every meeting of ticks 0..ticks-1 from the meeting log against the
    oracle's per-tick digests; returns the mismatches."""
    import hashlib
    bad = []
    log_rows = wb.meeting_log(start=0, sort=False)
    end = log_rows[:, 3] if len(log_rows) else np.zeros(0, dtype=np.uint32)
    for t in fx["ticks"][:ticks]:
        m = log_rows[end + 1 == t["tick"]]
        m = m[np.lexsort((m[:, 2], m[:, 1], m[:, 0]))] if len(m) else m
        if len(m) != t["meetings"] or \
                hashlib.sha256(np.ascontiguousarray(m, dtype=np.uint32).tobytes()).hexdigest() != t["meetings_sha256"]:
            bad.append(f"tick {t['tick']}: {len(m)} meetings, the oracle's fixture {t['meetings']}")
    return bad


def bench_serving(args) -> int:
    """Config 5: 20M resident objects, a seeded random 10% moving every tick
    (1 Hz), location updates ingested from a device update list; tracing
    queries arrive as a Poisson process at 50/s of simulated time (ticks are
    the simulated seconds), each with 1..10^4 seed pids and a window of
    60..3600 ticks (clipped to the history held by the run), 1 hop, or 2 hops
    for a fifth of them; queries are admitted by HBM footprint and overlap
    the ticks."""
    from glint_pipeline_amd import autotune, gen, runtime
    N = args.objects or 20_000_000
    H, K = args.history, args.steps
    w = gen.uniform_city(N, H + K + 1, gen.city_side(N), seed=7, grid_capacity=200)
    w.move_fraction = 0.1
    # the tuned parameters of the size class, with the FP32 guard-band test
    params = dict(autotune.params_for(autotune.size_class(N)), precision="fp32guard")
    if args.precision:
        params["precision"] = args.precision
    cfg = autotune.config_for(w, params, distinct_estimate("batched", N), graph_ticks=0,
                              meeting_log_capacity=1 << 29)
    g = gen.Generator(w)
    dw = runtime.DeviceWalkers(g)
    pids, upd, cnt = runtime.DeviceBuffer(4 * N), runtime.DeviceBuffer(16 * N), runtime.DeviceBuffer(4)
    full = runtime.DeviceBuffer(16 * N)
    full.upload(g.positions())
    rng = np.random.default_rng(11)
    wb = runtime.Workbench(cfg)
    sv = runtime.Server(wb, arena_bytes=8 << 30, max_running=16)
    t0 = time.perf_counter()
    moved = 0
    for t in range(H):
        if t == 0:
            sv.tick(0, d_full_xy=full.address())
        else:
            k = dw.step_updates(pids, upd, cnt)
            sv.tick(t, d_pids=pids.address(), d_xy=upd.address(), k=k)
        if t % 500 == 499:
            log(f"[bench] serving: history tick {t + 1}/{H}, {time.perf_counter() - t0:.0f} s, "
                f"log {wb.meeting_log_info()}, index {sv.index_info()}")
    log(f"[bench] serving: {N} objects, {H} ticks of history in {time.perf_counter() - t0:.1f} s; "
        f"log {wb.meeting_log_info()}; index {sv.index_info()}")
    tick_ms, ids = [], []
    pace = not args.no_pace

    def submit(t):
        seeds = rng.choice(N, size=int(rng.integers(1, 10_001)), replace=False)
        window = int(rng.integers(60, 3601))
        hops = 2 if rng.random() < 0.2 else 1
        ids.append(sv.submit(seeds, max(0, t - window), t, hops))

    t_start = time.perf_counter()
    for t in range(H, H + K):
        t1 = time.perf_counter()
        k = dw.step_updates(pids, upd, cnt)
        moved += k
        sv.tick(t, d_pids=pids.address(), d_xy=upd.address(), k=k)
        tick_ms.append((time.perf_counter() - t1) * 1e3)
        arrivals = np.sort(rng.uniform(0.0, 1.0, int(rng.poisson(50))))
        for a in arrivals:
            if pace:   # submit each query at its arrival time within the tick's second
                while time.perf_counter() < t1 + a:
                    sv.poll()
                    time.sleep(0.0005)
            submit(t)
            sv.poll()
        while pace and time.perf_counter() < t1 + 1.0:
            sv.poll()
            time.sleep(0.0005)
    elapsed_ticks = time.perf_counter() - t_start
    sv.drain()
    elapsed = time.perf_counter() - t_start
    qlat, hits, waited, retries = [], 0, 0, 0
    for qid in ids:
        info = sv.info(qid)
        qlat.append(info["done_ms"] - info["submit_ms"])
        waited += info["admitted_tick"] > info["submit_tick"]
        retries += info["retries"]
        h1, h2 = sv.result(qid)
        hits += len(h1) + len(h2)
    index_info = sv.index_info()
    wb_log = wb.meeting_log_info()
    sv.close()
    wb.close()
    for b in (pids, upd, cnt, full):
        b.free()
    lat = np.asarray(tick_ms)
    ql = np.asarray(qlat)
    out = {
        "metric": "object-updates/sec (moved objects) with concurrent tracing queries, 20M resident, 1 MI355X",
        # capacity: moved objects per second of tick processing (ingestion + full tick), queries running beside
        "value": round(moved / (float(np.sum(tick_ms)) / 1e3), 1), "unit": "object-updates/s", "n_gpus": 1,
        "steps": K, "warmup": H, "ms_per_step": round(float(np.mean(tick_ms)), 3),
        "paced": pace, "wall_seconds": round(elapsed_ticks, 2),
        "p50_tick_ms": round(float(np.percentile(lat, 50)), 3), "p99_tick_ms": round(float(np.percentile(lat, 99)), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded walk of resident objects, 10% moving per tick; Poisson queries)",
        "config": {"workload": f"serving: {N} resident objects, side {w.side:.0f} m, 10% moving per tick, "
                               f"queries Poisson 50/tick, seeds U{{1..10^4}}, windows U[60,3600] ticks "
                               f"({H} ticks of history before the first query), 20% two-hop",
                   "params": params, "arena_bytes": 8 << 30, "max_running": 16,
                   "meeting_log": {"capacity": 1 << 29, "logged": wb_log[0], "first_retained": wb_log[1]},
                   "meeting_index": index_info},
        "queries": {"count": len(ids), "p50_ms": round(float(np.percentile(ql, 50)), 3) if len(ql) else None,
                    "p99_ms": round(float(np.percentile(ql, 99)), 3) if len(ql) else None,
                    "admitted_later": int(waited), "retries": int(retries), "hits": int(hits),
                    "per_second": round(len(ids) / elapsed, 1)},
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_batched(args) -> int:
    """Config 4: 32 independent cities of uneven size in one batch (one graph
    replaying every city's ticks), against the same cities run one after the
    other."""
    from glint_pipeline_amd import autotune, gen, runtime
    W, K = args.warmup, args.steps
    G = 8 if args.graph_ticks is None else args.graph_ticks
    ws = gen.batched_cities(32, ticks=W + K)
    total = sum(w.num_objects for w in ws)
    t0 = time.perf_counter()
    if args.autotune:
        # one representative (the largest) city per size class; capacity is part of the result, not tuned
        table = autotune.load_table()
        for cls in ("small", "medium", "large"):
            members = [w for w in ws if autotune.size_class(w.num_objects) == cls]
            if not members:
                continue
            rep = max(members, key=lambda w: w.num_objects)
            res = autotune.tune(rep, tune_capacity=False, warmup=3, ticks=5, log=log)
            table[cls] = {"params": res.best, "ms_per_tick": round(res.best_ms, 3),
                          "workload": f"{rep.name}: {rep.num_objects} objects", "trials": res.trials}
        if args.save_tuned:
            autotune.save_table(table)
    bufs, cfgs, classes = [], [], []
    for w in ws:
        g = gen.Generator(w)
        b = runtime.DeviceBuffer(16 * w.num_objects * (W + K))
        b.upload(g.positions(), 0)
        runtime.DeviceWalkers(g).walk(W + K - 1, b, 16 * w.num_objects)
        bufs.append(b)
        cls = autotune.size_class(w.num_objects)
        p = dict(autotune.params_for(cls))
        if args.precision:
            p["precision"] = args.precision
        classes.append(cls)
        cfgs.append(autotune.config_for(w, p, distinct_estimate("batched", w.num_objects), graph_ticks=G))
    log(f"[bench] {len(ws)} cities, {total} objects, classes { {c: classes.count(c) for c in set(classes)} }; "
        f"traces generated in {time.perf_counter() - t0:.1f} s")

    # one batch: every city's ticks in one graph per G ticks
    bt = runtime.Batch(cfgs, graph_ticks=G)
    bt.run_ticks([b.address(0) for b in bufs], 0, W, latencies=False)
    t0 = time.perf_counter()
    ms = bt.run_ticks([b.address(16 * w.num_objects * W) for b, w in zip(bufs, ws)], W, K, latencies=True)
    elapsed = time.perf_counter() - t0
    meetings = sum(c.totals()["meetings"] for c in bt.contexts)
    bt.close()
    # the same cities one after the other, each with its own graph
    seq = 0.0
    for w, b, c in zip(ws, bufs, cfgs):
        with runtime.Workbench(c) as wb:
            wb.run_ticks(b.address(0), 0, W, latencies=False)
            wb.synchronize()
            t1 = time.perf_counter()
            wb.run_ticks(b.address(16 * w.num_objects * W), W, K, latencies=False)
            wb.synchronize()
            seq += time.perf_counter() - t1
    for b in bufs:
        b.free()
    lat = np.asarray(ms, dtype=np.float64).reshape(-1)   # every city's own ticks inside the batch
    out = {
        "metric": "object-updates/sec, 32 independent cities in one batch, 1 MI355X, exact contacts",
        "value": round(total * K / elapsed, 1), "unit": "object-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(elapsed * 1e3 / K, 3), "p50_tick_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_tick_ms": round(float(np.percentile(lat, 99)), 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded generator, positions resident in HBM)",
        "config": {"workload": f"batched uneven cities: 32 cities, N log-uniform in [50k, 2M] ({total} objects), "
                               f"side sqrt(N)*30 m, r in {{1.5,2,3}} m, min in {{3,5,10}}",
                   "graph_ticks": G, "meetings": meetings,
                   "sizes": [w.num_objects for w in ws]},
        "sequential_ms_per_tick": round(seq * 1e3 / K, 3),
        "batch_speedup": round(seq / elapsed, 3),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
