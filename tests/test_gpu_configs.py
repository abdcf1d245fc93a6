"""Three benchmark configurations at their full size against the oracle:

- the plain GPU form: 1M uniform walkers in a 30 km box, FP64, positions
  copied from host memory and the four stage calls every tick, 60 ticks,
  against the oracle run beside it;
- the batched cities: the 32 cities of 50k-2M objects of the batched
  benchmark in one batch (8-tick graphs), against committed oracle fixtures
  (tests/golden/make_config_golden.py) city by city and tick by tick;
- serving: 20M resident objects, a seeded 10% moving every tick, location
  updates ingested on the device, queries running beside the ticks; the
  served log against the committed fixture, and hop-1 queries against their
  definition on the served state."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from glint_pipeline_amd import autotune, gen, runtime
from tests.golden.make_config_golden import CITY_TICKS, SERVING_TICKS, cities, serving_workload
from tests.parity import device_trace, run_parity, split_log_by_tick

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).parent / "golden" / "configs_fixture.json"


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def check_against_fixture(wb, fx: dict, where: str) -> None:
    ticks = fx["ticks"]
    tot = wb.totals()
    assert tot["distinct_pairs"] == sum(t["contacts"] for t in ticks), f"{where}: distinct pairs"
    assert tot["tests"] == sum(t["tests"] for t in ticks), f"{where}: tested pairs"
    assert tot["meetings"] == sum(t["meetings"] for t in ticks), f"{where}: meeting counts"
    assert wb.meeting_log_info()[1] == 0
    log = split_log_by_tick(wb.meeting_log(sort=False))
    for t, ft in enumerate(ticks):
        m = log.get(t, np.zeros((0, 4), dtype=np.uint32))
        assert len(m) == ft["meetings"] and digest(m) == ft["meetings_sha256"], f"{where}: tick {t} meetings"
    k, s = wb.active_pairs()
    assert len(k) == fx["active_before_last"] and digest(k, s) == fx["active_before_last_sha256"], \
        f"{where}: active pairs"
    b, _, _ = wb.leaves()
    assert len(b) == fx["leaves"] and digest(b) == fx["leaves_sha256"], f"{where}: leaves"


def test_plain_gpu_form_1m_fp64_60_ticks():
    w = gen.plain_gpu_form()
    params = dict(autotune.DEFAULTS, precision="fp64")
    cfg = autotune.config_for(w, params, autotune.pairs_estimate("plain_gpu", w.num_objects), graph_ticks=0)
    tot = run_parity(w, 60, check_contacts=False, check_leaves_every=20, cfg=cfg)
    assert tot["meetings"] > 1000


@pytest.mark.skipif(not FIXTURE.exists(), reason="configuration fixtures not generated")
def test_batched_32_cities_match_fixtures():
    fx = json.loads(FIXTURE.read_text())
    ws = cities()
    assert [w.num_objects for w in ws] == [c["objects"] for c in fx["cities"]]
    cfgs = [autotune.config_for(w, autotune.params_for(autotune.size_class(w.num_objects)),
                                autotune.pairs_estimate("batched", w.num_objects), graph_ticks=8,
                                meeting_log_capacity=1 << 20) for w in ws]
    traces = [device_trace(w, CITY_TICKS) for w in ws]
    with runtime.Batch(cfgs, graph_ticks=8) as bt:
        bt.run_ticks([t.address(0) for t in traces], 0, 4, latencies=False)
        ms = bt.run_ticks([t.address(16 * w.num_objects * 4) for t, w in zip(traces, ws)], 4, CITY_TICKS - 4)
        assert (ms > 0).all()
        for i, (wb, c) in enumerate(zip(bt.contexts, fx["cities"])):
            check_against_fixture(wb, c, f"city {i} ({c['objects']} objects)")
    for t in traces:
        t.free()


@pytest.mark.skipif(not FIXTURE.exists(), reason="configuration fixtures not generated")
def test_serving_20m_resident_matches_fixture():
    fx = json.loads(FIXTURE.read_text())["serving"]
    w = serving_workload()
    N = w.num_objects
    # the serving bench's parameters: the size class's, with the FP32 guard band
    params = dict(autotune.params_for(autotune.size_class(N)), precision="fp32guard")
    cfg = autotune.config_for(w, params, autotune.pairs_estimate("batched", N), graph_ticks=0,
                              meeting_log_capacity=1 << 22)
    g = gen.Generator(w)
    dw = runtime.DeviceWalkers(g)
    pids, upd, cnt = runtime.DeviceBuffer(4 * N), runtime.DeviceBuffer(16 * N), runtime.DeviceBuffer(4)
    full = runtime.DeviceBuffer(16 * N)
    full.upload(g.positions())
    rng = np.random.default_rng(23)
    queries = []
    with runtime.Workbench(cfg) as wb, runtime.Server(wb, arena_bytes=1 << 30, max_running=4) as sv:
        moved = 0
        for t in range(SERVING_TICKS):
            if t == 0:
                sv.tick(0, d_full_xy=full.address())
            else:
                k = dw.step_updates(pids, upd, cnt)
                moved += k
                sv.tick(t, d_pids=pids.address(), d_xy=upd.address(), k=k)
            if t >= 6:   # queries beside the ticks, on the state of tick t
                keys, starts = wb.active_pairs()
                log_now = wb.meeting_log(sort=False)
                for _ in range(2):
                    seeds = rng.choice(N, size=2000, replace=False).astype(np.uint32)
                    queries.append((sv.submit(seeds, 0, t, 1), seeds, t, keys, starts, log_now))
                sv.poll()
        sv.drain()
        assert abs(moved / (SERVING_TICKS - 1) / N - 0.1) < 0.01
        check_against_fixture(wb, fx, "serving")
        for qid, seeds, t, keys, starts, log_now in queries:
            h1, h2 = sv.result(qid)
            # hop 1: meetings ended by t with a seed on either side, and pairs
            # in progress at t for at least min + 1 ticks, clipped to t
            a = (keys >> np.uint64(32)).astype(np.uint32)
            b = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            run = (t - starts.astype(np.int64)) >= w.min_meet_time + 1
            rows = [np.stack([log_now[:, 0], log_now[:, 1], log_now[:, 2], log_now[:, 3]], 1)[np.isin(log_now[:, 0], seeds)],
                    np.stack([log_now[:, 1], log_now[:, 0], log_now[:, 2], log_now[:, 3]], 1)[np.isin(log_now[:, 1], seeds)]]
            tt = np.full(len(keys), t, dtype=np.uint32)
            rows += [np.stack([a, b, starts, tt], 1)[run & np.isin(a, seeds)],
                     np.stack([b, a, starts, tt], 1)[run & np.isin(b, seeds)]]
            want = np.concatenate(rows).astype(np.uint32)
            want = want[np.lexsort((want[:, 2], want[:, 1], want[:, 0]))]
            assert np.array_equal(h1, want) and len(h2) == 0
    for buf in (pids, upd, cnt, full):
        buf.free()
