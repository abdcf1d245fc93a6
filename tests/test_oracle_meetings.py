"""The meeting rule (meeting.cpp:134-155) pinned by closed-form timelines and a
randomized model check (the pattern of the reference's hashtest: a table
against a plain model).

A pair in contact on the contiguous ticks s..e is emitted at tick e+1 as
(a, b, s, e) iff (e+1) - s >= min + 1, i.e. e - s >= min; a gap ends a run;
a run still active at the last tick is never emitted."""
import numpy as np
import pytest

from oracle.oracle import Oracle, pair_key


def run_timeline(seen: dict, ticks: int, min_meet: int, start_tick: int = 0):
    o = Oracle(16, (0, 0, 100, 100), 2.0, 4, min_meet_time=min_meet)
    out = []
    for t in range(start_tick, start_tick + ticks):
        keys = [pair_key(a, b) for (a, b), ts in seen.items() if t in ts]
        keys = keys + keys[:1]          # a duplicate discovery in the same tick
        o.evaluate_contacts(np.array(keys, dtype=np.uint64), t)
        out += [tuple(int(v) for v in m) for m in o.meetings()]
    o.close()
    return sorted(out)


def test_closed_form_timelines():
    seen = {
        (0, 1): set(range(0, 6)),                            # 6 ticks: 5 - 0 >= 5 -> emitted at 6
        (2, 3): set(range(0, 5)),                            # 5 ticks: 4 - 0 < 5 -> never
        (4, 5): set(range(0, 3)) | set(range(4, 11)),        # gap at 3; second run 4..10 emitted at 11
        (6, 7): set(range(3, 21)),                           # active at the last tick -> not emitted
        (1, 9): set(range(7, 13)) | set(range(14, 20)),      # two runs of 6: both emitted
    }
    got = run_timeline(seen, 21, 5)
    assert got == sorted([(0, 1, 0, 5), (4, 5, 4, 10), (1, 9, 7, 12), (1, 9, 14, 19)])


@pytest.mark.parametrize("min_meet", [0, 1, 3, 10])
def test_min_meet_threshold(min_meet):
    seen = {(0, 1): set(range(0, min_meet + 1)), (2, 3): set(range(0, min_meet))}
    got = run_timeline(seen, min_meet + 4, min_meet)
    # the shorter run (min_meet ticks) is one tick too short
    assert got == [(0, 1, 0, min_meet)]


def test_start_time_offset():
    seen = {(3, 8): set(range(100, 107))}
    assert run_timeline(seen, 10, 5, start_tick=100) == [(3, 8, 100, 106)]


def model_meetings(seen, ticks, min_meet, start=0):
    """Runs by a direct scan of each pair's tick set."""
    out = []
    last = start + ticks - 1
    for (a, b), ts in seen.items():
        t = start
        while t <= last:
            if t in ts:
                s = t
                while t + 1 in ts and t + 1 <= last:
                    t += 1
                e = t
                if e < last and e - s >= min_meet:
                    out.append((a, b, s, e))
            t += 1
    return sorted(out)


@pytest.mark.parametrize("seed", range(6))
def test_random_timelines_against_model(seed):
    rng = np.random.default_rng(seed)
    pairs = {}
    for _ in range(60):
        a, b = sorted(rng.choice(40, 2, replace=False).tolist())
        p = rng.uniform(0.3, 0.95)
        pairs[(a, b)] = {t for t in range(50) if rng.random() < p}
    mm = int(rng.integers(1, 8))
    assert run_timeline(pairs, 50, mm) == model_meetings(pairs, 50, mm)
