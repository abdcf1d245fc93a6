"""Summarise rocprofv3 CSV output (kernel stats, kernel trace, PMC counters)
as the text tables kept under profiles/.

    python -m glint_pipeline_amd.rocprof_summary <rocprofv3 output dir> [--kernel NAME]
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
from collections import defaultdict


def _short(name: str) -> str:
    n = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return n.split("(")[0].strip()


def _select(vals: list, skip: int, take: int | None) -> list:
    vals = vals[skip:]
    return vals[:take] if take else vals


def summarize(d: str, kernel: str | None = None, skip: int = 0, take: int | None = None) -> str:
    """skip/take select, per kernel and in dispatch order, the dispatches the
    trace and counter averages cover (e.g. the timed ticks of a run)."""
    out = []
    stats = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)
    for f in stats:
        out.append(f"# kernel stats ({os.path.relpath(f, d)})")
        out.append(f"{'kernel':28s} {'calls':>7s} {'total_ms':>11s} {'avg_ms':>10s} {'min_ms':>10s} {'max_ms':>10s} {'pct':>6s}")
        with open(f) as fh:
            rows = list(csv.DictReader(fh))
        for r in rows:
            name = _short(r.get("Name", r.get("KernelName", "?")))
            if kernel and kernel not in name:
                continue
            calls = int(float(r.get("Calls", 0)))
            tot = float(r.get("TotalDurationNs", 0)) / 1e6
            avg = float(r.get("AverageNs", 0)) / 1e6
            mn = float(r.get("MinNs", 0)) / 1e6
            mx = float(r.get("MaxNs", 0)) / 1e6
            pct = float(r.get("Percentage", 0))
            out.append(f"{name:28s} {calls:7d} {tot:11.3f} {avg:10.4f} {mn:10.4f} {mx:10.4f} {pct:6.2f}")
    if skip or take:
        for f in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
            per = defaultdict(list)
            with open(f) as fh:
                rows = sorted(csv.DictReader(fh), key=lambda r: int(r["Start_Timestamp"]))
            for r in rows:
                name = _short(r["Kernel_Name"])
                if kernel and kernel not in name:
                    continue
                per[name].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
            out.append(f"# kernel trace, dispatches {skip}..{skip + take - 1 if take else 'end'} of each kernel "
                       f"({os.path.relpath(f, d)})")
            out.append(f"{'kernel':28s} {'dispatches':>10s} {'avg_ms':>10s} {'min_ms':>10s} {'max_ms':>10s}")
            for name, vals in sorted(per.items(), key=lambda kv: -sum(kv[1])):
                v = _select(vals, skip, take)
                if v:
                    out.append(f"{name:28s} {len(v):10d} {sum(v) / len(v):10.4f} {min(v):10.4f} {max(v):10.4f}")
    pmc = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    for f in pmc:
        acc = defaultdict(lambda: defaultdict(list))
        with open(f) as fh:
            rows = list(csv.DictReader(fh))
        key = "Dispatch_Id" if rows and "Dispatch_Id" in rows[0] else None
        if key:
            rows.sort(key=lambda r: int(r[key]))
        for r in rows:
            name = _short(r.get("Kernel_Name", r.get("Kernel-Name", "?")))
            if kernel and kernel not in name:
                continue
            acc[name][r.get("Counter_Name", "?")].append(float(r.get("Counter_Value", 0)))
        sel = f", dispatches {skip}..{skip + take - 1 if take else 'end'} of each kernel" if skip or take else ""
        out.append(f"# counters per dispatch ({os.path.relpath(f, d)}{sel})")
        out.append(f"{'kernel':28s} {'counter':14s} {'dispatches':>10s} {'avg':>16s} {'min':>16s} {'max':>16s}")
        for name, cs in sorted(acc.items()):
            for cn, vals in sorted(cs.items()):
                v = _select(vals, skip, take)
                if not v:
                    continue
                out.append(f"{name:28s} {cn:14s} {len(v):10d} {sum(v) / len(v):16.1f} {min(v):16.1f} "
                           f"{max(v):16.1f}")
    if not out:
        out.append(f"no rocprofv3 CSV files under {d}")
    return "\n".join(out)


def _counter_avgs(d: str, counter: str, skip: int, take: int | None) -> dict:
    """{kernel: average of `counter` per dispatch} over the selected dispatches."""
    acc = defaultdict(list)
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            rows = list(csv.DictReader(fh))
        if rows and "Dispatch_Id" in rows[0]:
            rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        for r in rows:
            if r.get("Counter_Name") == counter:
                acc[_short(r.get("Kernel_Name", "?"))].append(float(r.get("Counter_Value", 0)))
    return {k: sum(v) / len(v) for k, v in ((k, _select(v, skip, take)) for k, v in acc.items()) if v}


def traffic(fetch_dir: str, write_dir: str, copy_fetch_dir: str, copy_write_dir: str, copy_bytes: int,
            skip: int, take: int | None) -> str:
    """HBM bytes per launch of every kernel: FETCH_SIZE and WRITE_SIZE (KiB
    per dispatch) scaled by what the same counters read for a plain copy of
    copy_bytes bytes (which reads and writes copy_bytes each).  The last
    lines are machine-readable: TRAFFIC_BYTES_PER_LAUNCH <kernel> <bytes>
    and SOURCE_SHA <digest of the kernel sources measured>."""
    from .build import source_digest
    fetch = _counter_avgs(fetch_dir, "FETCH_SIZE", skip, take)
    write = _counter_avgs(write_dir, "WRITE_SIZE", skip, take)
    cf = _counter_avgs(copy_fetch_dir, "FETCH_SIZE", 0, None).get("k_copy")
    cw = _counter_avgs(copy_write_dir, "WRITE_SIZE", 0, None).get("k_copy")
    if not cf or not cw:
        raise SystemExit("no k_copy calibration counters found")
    ff, fw = copy_bytes / 1024.0 / cf, copy_bytes / 1024.0 / cw
    out = [f"# calibration: a {copy_bytes} B copy reads FETCH_SIZE {cf:.1f} KiB and WRITE_SIZE {cw:.1f} KiB:",
           f"# bytes = FETCH_SIZE * {ff:.3f} * 1024 + WRITE_SIZE * {fw:.3f} * 1024",
           f"{'kernel':28s} {'FETCH_KiB':>14s} {'WRITE_KiB':>14s} {'read_GB':>10s} {'written_GB':>10s} {'total_GB':>10s}"]
    base = defaultdict(float)   # template variants of one kernel summed (each runs once a tick)
    for k in sorted(set(fetch) | set(write), key=lambda k: -(fetch.get(k, 0) * ff + write.get(k, 0) * fw)):
        rb, wb = fetch.get(k, 0.0) * ff * 1024, write.get(k, 0.0) * fw * 1024
        out.append(f"{k:28s} {fetch.get(k, 0):14.1f} {write.get(k, 0):14.1f} {rb / 1e9:10.4f} {wb / 1e9:10.4f} "
                   f"{(rb + wb) / 1e9:10.4f}")
        base[k.split('<')[0]] += rb + wb
    lines = [f"TRAFFIC_BYTES_PER_LAUNCH {k} {v:.0f}" for k, v in sorted(base.items(), key=lambda kv: -kv[1])]
    return "\n".join(out + [""] + lines + [f"SOURCE_SHA {source_digest()}"])


def valu(d: str, skip: int, take: int | None) -> str:
    """The counter table of a run with SQ_INSTS_VALU among its counters, then
    machine-readable lines: VALU_INSTS_PER_LAUNCH <kernel> <wave instructions>
    (template variants summed) and SOURCE_SHA <digest of the kernel sources>."""
    from .build import source_digest
    insts = _counter_avgs(d, "SQ_INSTS_VALU", skip, take)
    names = ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE", "SQ_WAVES"]
    cols = {n: _counter_avgs(d, n, skip, take) for n in names}
    sel = f"dispatches {skip}..{skip + take - 1 if take else 'end'} of each kernel"
    out = [f"# counters per dispatch, averages over {sel} (kernels of at least 10^5 VALU instructions)",
           f"{'kernel':28s} " + " ".join(f"{n:>20s}" for n in names)]
    for k, v in sorted(insts.items(), key=lambda kv: -kv[1]):
        if v >= 1e5:
            out.append(f"{k:28s} " + " ".join(f"{cols[n].get(k, 0.0):20.1f}" for n in names))
    base = defaultdict(float)
    for k, v in insts.items():
        base[k.split('<')[0]] += v
    lines = [f"VALU_INSTS_PER_LAUNCH {k} {v:.0f}" for k, v in sorted(base.items(), key=lambda kv: -kv[1]) if v >= 1e5]
    return "\n".join(out + [""] + lines + [f"SOURCE_SHA {source_digest()}"])


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--kernel", default=None)
    ap.add_argument("--skip", type=int, default=0, help="dispatches of each kernel to leave out first")
    ap.add_argument("--take", type=int, default=None, help="dispatches of each kernel to average after those")
    ap.add_argument("--traffic", nargs=4, metavar=("WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR", "COPY_BYTES"),
                    help="dir holds a FETCH_SIZE run: per-kernel HBM bytes per launch, calibrated by a copy")
    ap.add_argument("--valu", action="store_true", help="dir holds an SQ_INSTS_VALU run: VALU instructions per launch")
    a = ap.parse_args()
    if a.valu:
        print(valu(a.dir, a.skip, a.take))
    elif a.traffic:
        print(traffic(a.dir, a.traffic[0], a.traffic[1], a.traffic[2], int(a.traffic[3]), a.skip, a.take))
    else:
        print(summarize(a.dir, a.kernel, a.skip, a.take))


if __name__ == "__main__":
    main()
