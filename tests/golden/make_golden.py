"""Writes tests/golden/oracle_fixtures.json: for small seeded workloads, the
oracle's per-tick contact and meeting counts, a SHA-256 of every tick's sorted
meeting list and of the final leaf boxes.  The fixtures pin the oracle against
regressions and give the GPU path a checker that does not run the oracle.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from glint_pipeline_amd import gen  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

CASES = [
    dict(name="uniform_2k_cap16", n=2000, side=500.0, seed=101, cap=16, r=2.0, mm=5, ticks=50),
    dict(name="uniform_10k_cap64_r3_min3", n=10000, side=1000.0, seed=102, cap=64, r=3.0, mm=3, ticks=50),
    dict(name="hotspot_10k_cap200_r1p5_min10", n=10000, side=3000.0, seed=103, cap=200, r=1.5, mm=10, ticks=50,
         spots=8, sigma=80.0),
    dict(name="uniform_500_cap16", n=500, side=200.0, seed=104, cap=16, r=2.0, mm=5, ticks=50),
    dict(name="hotspot_20k_cap64_min5", n=20000, side=2000.0, seed=105, cap=64, r=2.0, mm=5, ticks=40,
         spots=16, sigma=60.0),
]


def workload(c) -> gen.Workload:
    return gen.Workload(c["name"], c["n"], c["ticks"], c["side"], seed=c["seed"], num_hotspots=c.get("spots", 0),
                        hotspot_fraction=0.8 if c.get("spots") else 0.0, hotspot_sigma=c.get("sigma", 200.0),
                        reach_distance=c["r"], min_meet_time=c["mm"], grid_capacity=c["cap"])


def run(c) -> dict:
    w = workload(c)
    g = gen.Generator(w)
    o = Oracle(w.num_objects, w.box(), w.reach_distance, w.grid_capacity, w.min_meet_time)
    h = hashlib.sha256()
    contacts, meetings, sample = [], [], []
    for t in range(w.ticks):
        o.tick(g.positions() if t == 0 else g.step(), t)
        m = o.meetings()
        h.update(np.ascontiguousarray(m, dtype=np.uint32).tobytes())
        contacts.append(int(o.stats()["num_contacts"]))
        meetings.append(int(len(m)))
        if len(sample) < 12:
            sample += [[int(v) for v in row] for row in m[: 12 - len(sample)]]
    boxes, levels, counts = o.leaves()
    o.close()
    return {"case": c, "contacts_per_tick": contacts, "meetings_per_tick": meetings,
            "meetings_sha256": h.hexdigest(), "meetings_sample": sample,
            "leaves": int(len(boxes)),
            "leaves_sha256": hashlib.sha256(np.ascontiguousarray(boxes).tobytes()).hexdigest()}


def main() -> None:
    out = [run(c) for c in CASES]
    (Path(__file__).parent / "oracle_fixtures.json").write_text(json.dumps(out, indent=1) + "\n")
    for r in out:
        print(r["case"]["name"], sum(r["contacts_per_tick"]), sum(r["meetings_per_tick"]), r["leaves"])


if __name__ == "__main__":
    main()
