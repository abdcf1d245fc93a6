"""Shared driver of the GPU-vs-oracle parity tests: the same seeded positions
go to the HIP path (through the C ABI) and to the CPU oracle, tick by tick."""
from __future__ import annotations

import numpy as np

from glint_pipeline_amd import gen, runtime
from oracle.oracle import Oracle


def keys_of(meetings: np.ndarray) -> np.ndarray:
    return (meetings[:, 0].astype(np.uint64) << np.uint64(32)) | meetings[:, 1].astype(np.uint64)


def run_parity(w: gen.Workload, ticks: int, *, check_contacts: bool = True, check_leaves_every: int = 10,
               complete: bool = False, precision: int = runtime.PRECISION_FP64,
               dynamic_schema: int = runtime.SCHEMA_DYNAMIC, box=None, **cfg_kw) -> dict:
    """Runs both paths for `ticks` ticks and asserts identical per-tick contact
    sets, meeting sets and leaf boxes.  Returns totals for the caller."""
    g = gen.Generator(w)
    box = box or w.box()
    o = Oracle(w.num_objects, box, w.reach_distance, w.grid_capacity, w.min_meet_time, complete=complete,
               dynamic_schema=dynamic_schema)
    kw = dict(reach_distance=w.reach_distance, grid_capacity=w.grid_capacity, min_meet_time=w.min_meet_time,
              complete=complete, precision=precision, record_contacts=check_contacts,
              dynamic_schema=dynamic_schema)
    kw.update(cfg_kw)
    cfg = runtime.make_config(w.num_objects, box, **kw)
    totals = {"meetings": 0, "contacts": 0}
    with runtime.Workbench(cfg) as wb:
        for t in range(ticks):
            xy = g.positions() if t == 0 else g.step()
            o.tick(xy, t)
            m = wb.tick(xy, t)
            om = o.meetings()
            assert m.shape == om.shape and np.array_equal(m, om), (
                f"tick {t}: meetings differ ({len(m)} GPU vs {len(om)} oracle)")
            totals["meetings"] += len(m)
            if check_contacts:
                c = wb.contacts()
                oc = o.contacts()
                assert np.array_equal(c, oc), f"tick {t}: contacts differ ({len(c)} GPU vs {len(oc)} oracle)"
                totals["contacts"] += len(c)
            st = wb.stats()
            assert st["num_active"] == o.stats()["num_active"], f"tick {t}: active pair counts differ"
            if check_leaves_every and t % check_leaves_every == 0:
                b, lv, cnt = wb.leaves()
                ob, olv, ocnt = o.leaves()
                assert np.array_equal(b, ob) and np.array_equal(lv, olv), f"tick {t}: leaf boxes differ"
        k, s = wb.active_pairs()
        ok, os_ = o.active()
        if not check_contacts and w.min_meet_time != 0:
            # pairs first seen at the last tick are not kept without recorded contacts
            keep = os_ != ticks - 1
            ok, os_ = ok[keep], os_[keep]
        assert np.array_equal(k, ok) and np.array_equal(s, os_), "active pairs after the last tick differ"
    o.close()
    return totals
