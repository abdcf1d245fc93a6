"""Writes tests/golden/configs_fixture.json: the oracle's results on two
benchmark configurations at their full size, for the GPU tests to check
without running the oracle on the GPU box.

- batched cities: the 32 cities of the batched benchmark (N log-uniform in
  [50k, 2M], side sqrt(N)*30 m, reach in {1.5, 2, 3} m, minimum meeting in
  {3, 5, 10}), 12 ticks each;
- serving: 20M resident objects of the serving benchmark, a seeded 10% of
  them moving every tick, 8 ticks.

For every city and tick: unique contacts, candidate pairs tested, meetings
(count and SHA-256 of the (pid_a, pid_b, start, end) rows sorted by pid_a,
pid_b, start).  After the last tick: the active pairs whose run began before
it (SHA-256 of the sorted keys then the starts) and the leaf boxes.

    python tests/golden/make_config_golden.py
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from glint_pipeline_amd import gen  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

OUT = Path(__file__).parent / "configs_fixture.json"
CITY_TICKS = 12
SERVING_OBJECTS = 20_000_000
SERVING_TICKS = 8


def cities() -> list[gen.Workload]:
    return gen.batched_cities(32, ticks=CITY_TICKS)


def serving_workload() -> gen.Workload:
    w = gen.uniform_city(SERVING_OBJECTS, SERVING_TICKS, gen.city_side(SERVING_OBJECTS), seed=7, grid_capacity=200)
    w.move_fraction = 0.1
    return w


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def run(w: gen.Workload, ticks: int) -> dict:
    g = gen.Generator(w)
    o = Oracle(w.num_objects, w.box(), w.reach_distance, w.grid_capacity, w.min_meet_time)
    per = []
    for t in range(ticks):
        o.tick(g.positions() if t == 0 else g.step(), t)
        st = o.stats()
        m = o.meetings()
        per.append({"contacts": int(st["num_contacts"]), "tests": int(st["num_tests"]), "meetings": int(len(m)),
                    "meetings_sha256": digest(m.astype(np.uint32))})
    keys, starts = o.active()
    keep = starts != ticks - 1
    boxes, _, _ = o.leaves()
    out = {"name": w.name, "objects": w.num_objects, "ticks": per,
           "active_before_last": int(keep.sum()),
           "active_before_last_sha256": digest(keys[keep].astype(np.uint64), starts[keep].astype(np.uint32)),
           "leaves": int(len(boxes)), "leaves_sha256": digest(boxes.astype(np.float64))}
    o.close()
    return out


def main() -> None:
    t0 = time.perf_counter()
    bc = []
    for w in cities():
        bc.append(run(w, CITY_TICKS))
        print(f"{w.name}: {w.num_objects} objects, {sum(t['meetings'] for t in bc[-1]['ticks'])} meetings, "
              f"{time.perf_counter() - t0:.0f} s", flush=True)
    sv = run(serving_workload(), SERVING_TICKS)
    print(f"serving: {sum(t['meetings'] for t in sv['ticks'])} meetings, {time.perf_counter() - t0:.0f} s", flush=True)
    OUT.write_text(json.dumps({"city_ticks": CITY_TICKS, "cities": bc, "serving_ticks": SERVING_TICKS,
                               "serving": sv}, indent=1) + "\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
