"""Committed fixtures (tests/golden/oracle_fixtures.json, written by
tests/golden/make_golden.py) against the oracle (CPU) and against the HIP path
(GPU, without running the oracle)."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime
from tests.golden.make_golden import run, workload

FIX = json.loads((Path(__file__).parent / "golden" / "oracle_fixtures.json").read_text())


@pytest.mark.parametrize("fx", FIX, ids=[f["case"]["name"] for f in FIX])
def test_oracle_reproduces_fixture(fx):
    assert run(fx["case"]) == fx


@pytest.mark.gpu
@pytest.mark.parametrize("fx", FIX, ids=[f["case"]["name"] for f in FIX])
def test_gpu_reproduces_fixture(fx):
    w = workload(fx["case"])
    g = gen.Generator(w)
    cfg = runtime.make_config(w.num_objects, w.box(), reach_distance=w.reach_distance,
                              grid_capacity=w.grid_capacity, min_meet_time=w.min_meet_time, record_contacts=True)
    h = hashlib.sha256()
    with runtime.Workbench(cfg) as wb:
        for t in range(w.ticks):
            m = wb.tick(g.positions() if t == 0 else g.step(), t)
            h.update(np.ascontiguousarray(m, dtype=np.uint32).tobytes())
            assert len(wb.contacts()) == fx["contacts_per_tick"][t]
            assert len(m) == fx["meetings_per_tick"][t]
        boxes, _, _ = wb.leaves()
    assert h.hexdigest() == fx["meetings_sha256"]
    assert hashlib.sha256(np.ascontiguousarray(boxes).tobytes()).hexdigest() == fx["leaves_sha256"]
