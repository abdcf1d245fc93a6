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


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--kernel", default=None)
    ap.add_argument("--skip", type=int, default=0, help="dispatches of each kernel to leave out first")
    ap.add_argument("--take", type=int, default=None, help="dispatches of each kernel to average after those")
    a = ap.parse_args()
    print(summarize(a.dir, a.kernel, a.skip, a.take))


if __name__ == "__main__":
    main()
