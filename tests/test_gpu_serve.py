"""The serving loop: location updates (a random 10% subset moves each tick)
ingested into resident state, and tracing queries over windows answered from
the meeting log and the active table while ticks continue.

- the log of the served run equals the oracle's meetings on the same walk;
- every query equals a host restatement of its definition (hop 1: meetings
  ended by the submission tick T, plus pairs in progress at T with
  T - start >= min + 1, clipped to T; hop 2: meetings ended by T of the
  hop-1 contacts after their first exposure);
- each query gives the same hits alone and in a mix where a small arena
  makes queries wait (admitted ticks later) or overflow and retry."""
import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime

pytestmark = pytest.mark.gpu

N, SIDE, TICKS = 8000, 1000.0, 40


def workload():
    w = gen.uniform_city(N, TICKS, SIDE, seed=17, grid_capacity=32, min_meet_time=3)
    w.move_fraction = 0.1
    return w


def plan(seed=5):
    """(tick, seeds, t_start, t_end, hops) queries submitted during the run."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(8, TICKS, 3):
        for _ in range(int(rng.integers(1, 4))):
            seeds = rng.choice(N, size=int(rng.integers(1, 300)), replace=False)
            w = int(rng.integers(3, t + 1))
            out.append((t, seeds, t - w, t, int(rng.integers(1, 3))))
    return out


def serve(queries, arena, max_running, min_hits=4096, alone=False, poll=True, estimate_percent=100,
          index_block_ticks=4):
    """Runs the walk under the server; returns per-query (hop1, hop2, info),
    the final log, the log totals after each tick, and active pairs at every
    submission tick."""
    w = workload()
    g = gen.Generator(w)
    dw = runtime.DeviceWalkers(g)
    cfg = runtime.make_config(N, w.box(), reach_distance=w.reach_distance, grid_capacity=w.grid_capacity,
                              min_meet_time=w.min_meet_time, meeting_log_capacity=1 << 20)
    pids = runtime.DeviceBuffer(4 * N)
    upd = runtime.DeviceBuffer(16 * N)
    cnt = runtime.DeviceBuffer(4)
    full = runtime.DeviceBuffer(16 * N)
    full.upload(g.positions())
    results, totals, actives = {}, [], {}
    with runtime.Workbench(cfg) as wb, runtime.Server(wb, arena_bytes=arena, max_running=max_running,
                                                      min_hits=min_hits, estimate_percent=estimate_percent,
                                                      index_block_ticks=index_block_ticks) as sv:
        ids = {}
        for t in range(TICKS):
            if t == 0:
                sv.tick(0, d_full_xy=full.address())
            else:
                k = dw.step_updates(pids, upd, cnt)
                sv.tick(t, d_pids=pids.address(), d_xy=upd.address(), k=k)
            totals.append(wb.meeting_log_info()[0])
            todo = [(i, q) for i, q in enumerate(queries) if q[0] == t]
            if todo:
                actives[t] = wb.active_pairs()
            for i, (_, seeds, t0, t1, hops) in todo:
                ids[i] = sv.submit(seeds, t0, t1, hops)
                if alone:
                    sv.drain()
            if poll:
                sv.poll()
        sv.drain()
        info = sv.index_info()
        if index_block_ticks != 0xFFFFFFFF:
            assert info["blocks"] >= TICKS // index_block_ticks - 1 and info["entries"] == 2 * info["sealed"]
        else:
            assert info["blocks"] == 0
        for i, qid in ids.items():
            info = sv.info(qid)
            h1, h2 = sv.result(qid)
            results[i] = (h1, h2, info)
        log = wb.meeting_log(sort=False)
    for b in (pids, upd, cnt, full):
        b.free()
    return results, log, totals, actives


def expected(q, log, totals, actives, min_meet):
    t, seeds, t0, t1, hops = q
    S = set(int(v) for v in seeds)
    ended = log[:totals[t]].astype(np.int64)
    h1 = []
    for a, b, s, e in ended:
        if s <= t1 and e >= t0:
            if a in S:
                h1.append((a, b, s, e))
            if b in S:
                h1.append((b, a, s, e))
    keys, starts = actives[t]
    for k, s in zip(keys.astype(np.uint64), starts.astype(np.int64)):
        a, b = int(k >> np.uint64(32)), int(k & np.uint64(0xFFFFFFFF))
        if s <= t and t - s >= min_meet + 1 and s <= t1 and t >= t0:
            if a in S:
                h1.append((a, b, s, t))
            if b in S:
                h1.append((b, a, s, t))
    h2 = []
    if hops == 2:
        first = {}
        for _, c, s, _ in h1:
            if c not in S:
                first[c] = min(first.get(c, s), s)
        for a, b, s, e in ended:
            if s <= t1 and e >= t0:
                if a in first and e >= first[a]:
                    h2.append((a, b, s, e))
                if b in first and e >= first[b]:
                    h2.append((b, a, s, e))
    srt = lambda v: np.array(sorted(v), dtype=np.uint32).reshape(-1, 4)
    return srt(h1), srt(h2)


def test_served_log_matches_oracle():
    from oracle.oracle import Oracle
    w = workload()
    g = gen.Generator(w)
    o = Oracle(N, w.box(), w.reach_distance, w.grid_capacity, w.min_meet_time)
    want = []
    for t in range(TICKS):
        o.tick(g.positions() if t == 0 else g.step(), t)
        want.append(o.meetings())
    o.close()
    want = np.concatenate(want)
    want = want[np.lexsort((want[:, 3], want[:, 2], want[:, 1], want[:, 0]))]
    _, log, _, _ = serve([], 1 << 26, 2)
    got = log[np.lexsort((log[:, 3], log[:, 2], log[:, 1], log[:, 0]))]
    assert len(want) > 100 and np.array_equal(got, want)


@pytest.mark.parametrize("index_block_ticks", [4, 7, 0xFFFFFFFF])
def test_queries_alone_and_in_a_mix(index_block_ticks):
    """Blocks of 4 or 7 ticks of the per-object meeting index searched with
    the log tail scanned, or the whole log scanned (no index): the same
    answers, each its definition."""
    queries = plan()
    print(f"[serve] {len(queries)} queries, index blocks of {index_block_ticks} ticks", flush=True)
    alone, log, totals, actives = serve(queries, 1 << 28, 1, alone=True, index_block_ticks=index_block_ticks)
    # a 256 KB arena and two slots, admissions only at ticks: queries wait
    # for later ticks, and tiny hit buffers force re-admission after overflow
    mix, log2, _, _ = serve(queries, 256 << 10, 2, min_hits=16, poll=False, estimate_percent=5,
                            index_block_ticks=index_block_ticks)
    srt = lambda v: v[np.lexsort((v[:, 3], v[:, 2], v[:, 1], v[:, 0]))]
    assert np.array_equal(srt(log), srt(log2))
    waited = retried = 0
    for i, q in enumerate(queries):
        w1, w2 = expected(q, log, totals, actives, 3)
        a1, a2, ai = alone[i]
        m1, m2, mi = mix[i]
        assert np.array_equal(a1, w1) and np.array_equal(a2, w2), f"query {i} alone differs from its definition"
        assert np.array_equal(m1, a1) and np.array_equal(m2, a2), f"query {i}: mix differs from alone"
        assert ai["state"] == mi["state"] == runtime.QUERY_DONE
        waited += mi["admitted_tick"] > mi["submit_tick"]
        retried += mi["retries"] > 0
    assert sum(len(alone[i][0]) for i in alone) > 100
    assert sum(len(alone[i][1]) for i in alone) > 100
    assert waited > 0 and retried > 0
