"""glint_run_ticks_noalloc (per-tick stamps in the library's fixed device
array, plain launches) against glint_run_ticks on the same inputs, through
the C ABI: the same meetings tick by tick, totals, active pairs and leaves,
on two calls in a row, the second shorter than the first."""
import ctypes

import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime
from tests.parity import device_trace, split_log_by_tick

pytestmark = pytest.mark.gpu


def _run(wb, fn, trace, n, first, nticks):
    ms = np.zeros(nticks, dtype=np.float32)
    st = fn(wb.h, ctypes.c_void_p(trace.address(16 * n * first)), first, nticks, ms.ctypes.data)
    assert st == runtime.GLINT_OK, runtime.lib().glint_last_error_message(wb.h)
    return ms


def test_noalloc_run_ticks_equals_run_ticks():
    w = gen.Workload("hot_noalloc", 30_000, 20, 3000.0, seed=81, num_hotspots=8, hotspot_fraction=0.8,
                     hotspot_sigma=60.0, grid_capacity=64)
    trace = device_trace(w, w.ticks)
    cfg = runtime.make_config(w.num_objects, w.box(), grid_capacity=w.grid_capacity,
                              precision=runtime.PRECISION_FP32_GUARD, graph_ticks=8, meeting_log_capacity=1 << 22)
    L = runtime.lib()
    with runtime.Workbench(cfg) as a, runtime.Workbench(cfg) as b:
        for first, nticks in [(0, 13), (13, 7)]:    # the second call is shorter
            ma = _run(a, L.glint_run_ticks, trace, w.num_objects, first, nticks)
            mb = _run(b, L.glint_run_ticks_noalloc, trace, w.num_objects, first, nticks)
            assert (ma > 0).all() and (mb > 0).all()
            # every total but the pair slots evaluated, a work count that
            # depends on the order objects land in the join's tiles
            ta, tb = a.totals(), b.totals()
            ta.pop("evaluated")
            tb.pop("evaluated")
            assert ta == tb
            assert a.stats()["num_active"] == b.stats()["num_active"]
            la = split_log_by_tick(a.meeting_log(sort=False))
            lb = split_log_by_tick(b.meeting_log(sort=False))
            assert sorted(la) == sorted(lb) and all(np.array_equal(la[t], lb[t]) for t in la)
        assert ta["meetings"] > 1000
        ka, sa = a.active_pairs()
        kb, sb = b.active_pairs()
        assert np.array_equal(ka, kb) and np.array_equal(sa, sb)
        ba, va, ca = a.leaves()
        bb, vb, cb = b.leaves()
        assert np.array_equal(ba, bb) and np.array_equal(va, vb) and np.array_equal(ca, cb)
        # more ticks than its fixed array holds: refused before anything runs
        ms = np.zeros(4097, dtype=np.float32)
        st = L.glint_run_ticks_noalloc(b.h, ctypes.c_void_p(trace.address(0)), 20, 4097, ms.ctypes.data)
        assert st == runtime.GLINT_ERR_INVALID_ARGUMENT
        assert b"4096" in L.glint_last_error_message(b.h)
    trace.free()
