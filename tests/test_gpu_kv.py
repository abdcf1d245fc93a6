"""Sorted key-value output on the device against the oracle, bit for bit:
every part of every flush (sorted keys and boxes of both buffers, sids,
file records, bitmaps, the packed key stream) on the same seeded traces."""
import numpy as np
import pytest

from glint_pipeline_amd import runtime
from tests.kvcases import KvCase, run_device, run_oracle

pytestmark = pytest.mark.gpu

CASES = {
    "plain": KvCase(max_meet=4096, area_limit=1e9),
    "cut": KvCase(max_meet=9, area_limit=600.0, seed=7),
    "wide": KvCase(n=6000, side=1200.0, area_limit=3000.0, seed=11),
    "hotspots": KvCase(n=5000, side=800.0, hotspots=4, split=10, area_limit=5000.0, seed=13, flush_at=(20, 44, 59)),
}


def assert_flush_equal(got: dict, want: dict, where: str):
    assert got["summary"] == want["summary"], where
    for part in ("keys", "okeys", "sids", "bitmaps", "obitmap", "packed"):
        assert np.array_equal(got[part], want[part]), f"{where}: {part}"
    for part in ("boxes", "oboxes"):
        assert np.array_equal(got[part].view(np.uint32), want[part].view(np.uint32)), f"{where}: {part}"
    g, w = got["ctfs"], want["ctfs"]
    assert len(g) == len(w)
    for name in w.dtype.names:
        a, b = g[name], w[name]
        if a.dtype.kind == "f":
            a, b = a.view(np.uint32), b.view(np.uint32)
        assert np.array_equal(a, b), f"{where}: ctf field {name}"


@pytest.mark.parametrize("name", sorted(CASES))
def test_kv_flushes_match_oracle(name):
    case = CASES[name]
    print(f"[kv] case {name} n={case.n} side={case.side}", flush=True)
    want = run_oracle(case)
    got = run_device(case)
    for t, (a, b) in enumerate(zip(got.meetings, want.meetings)):
        assert np.array_equal(a, b), f"{name}: meetings at tick {t}"
    assert len(got.flushes) == len(want.flushes) == len(case.flush_at)
    for k, (g, w) in enumerate(zip(got.flushes, want.flushes)):
        assert w["summary"]["kept_keys"] > 0
        assert_flush_equal(g, w, f"{name} flush {k}")


def test_kv_with_graph_batches_matches_oracle():
    case = KvCase(max_meet=9, area_limit=600.0, seed=7, flush_at=(31, 59))
    print("[kv] graph batches", flush=True)
    want = run_oracle(case)
    got = run_device(case, graph_ticks=4)
    for k, (g, w) in enumerate(zip(got.flushes, want.flushes)):
        assert_flush_equal(g, w, f"graph flush {k}")


def test_kv_empty_flush_and_off_state():
    case = KvCase(ticks=3, flush_at=(0, 2))
    want = run_oracle(case)
    got = run_device(case)
    for k, (g, w) in enumerate(zip(got.flushes, want.flushes)):
        assert g["summary"]["kept_keys"] == 0
        assert_flush_equal(g, w, f"empty flush {k}")
    cfg = runtime.make_config(100, (0, 0, 100, 100))
    with runtime.Workbench(cfg) as wb:
        with pytest.raises(runtime.GlintError) as e:
            wb.kv_flush()
        assert e.value.status == runtime.GLINT_ERR_STATE


def test_kv_log_overflow_is_an_error():
    case = KvCase(ticks=40, flush_at=(39,))
    with pytest.raises(runtime.GlintError) as e:
        run_device(case, kv_capacity=8)
    assert e.value.status == runtime.GLINT_ERR_OUTPUT_CAPACITY
