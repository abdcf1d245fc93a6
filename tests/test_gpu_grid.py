"""The per-tick grid rebuild (GLINT_SCHEMA_GRID_REBUILD, the reference's GPU
path with dynamic_schema off) on the GPU against the oracle: the same tree
every tick (leaf boxes and depths bit-equal), the same contacts, meetings
and active pairs, with and without recorded contacts, in FP64 and with the
FP32 guard band, in plain launches and in captured graphs; the cell wrap on
the high x edge and the refused high y edge."""
import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime
from tests.parity import run_parity, run_ticks_parity

pytestmark = pytest.mark.gpu

GRID = runtime.SCHEMA_GRID_REBUILD


def margin_box(w, m=1.0):
    # the generator's walkers reach the box edge; the grid has no cell on the high y edge
    return (-m, -m, w.side + m, w.side + m)


def test_grid_rebuild_parity_uniform():
    w = gen.uniform_city(3000, 40, 500.0, seed=11, grid_capacity=16)
    tot = run_parity(w, 40, dynamic_schema=GRID, box=margin_box(w), check_leaves_every=1)
    assert tot["meetings"] > 20 and tot["contacts"] > 1000


@pytest.mark.parametrize("check_contacts", [True, False])
def test_grid_rebuild_parity_hotspots_deep_trees(check_contacts):
    """Depths 3 to 14: the fine histograms of coarse blocks over capacity."""
    w = gen.Workload("hot", 30_000, 12, 8000.0, seed=5, num_hotspots=8, hotspot_fraction=0.9,
                     hotspot_sigma=40.0, grid_capacity=32)
    run_parity(w, 12, dynamic_schema=GRID, box=margin_box(w), check_contacts=check_contacts,
               check_leaves_every=3, num_meeting_buckets=1 << 21, contact_capacity=4_000_000)


@pytest.mark.parametrize("check_contacts", [True, False])
def test_grid_rebuild_parity_cells_over_capacity_guard_band(check_contacts):
    """Clusters of sigma 3 m: cells over capacity get four leaves of depth 15,
    and lookups there accept many leaves (candidate lists that overflow)."""
    w = gen.Workload("clusters", 30_000, 8, 30_000.0, seed=5, num_hotspots=4, hotspot_fraction=0.05,
                     hotspot_sigma=3.0, speed_min=0.1, speed_max=0.5, grid_capacity=8)
    run_parity(w, 8, dynamic_schema=GRID, box=margin_box(w), precision=runtime.PRECISION_FP32_GUARD,
               check_leaves_every=2, join_tile=256, num_meeting_buckets=1 << 20, check_contacts=check_contacts)


@pytest.mark.parametrize("graph_ticks", [0, 4])
def test_grid_rebuild_run_ticks_matches_oracle(graph_ticks):
    """Ticks from a device trace, with and without graphs: every tick's
    meetings (through the meeting log), active pairs, leaf boxes, depths and
    populations, and the node count equal the oracle's."""
    w = gen.Workload("hot", 20_000, 20, 2000.0, seed=31, num_hotspots=6, hotspot_fraction=0.7,
                     hotspot_sigma=30.0, grid_capacity=24)

    def extra(wb, o):
        b, lv, cnt = wb.leaves()
        ob, olv, ocnt = o.leaves()
        assert np.array_equal(cnt, ocnt)
        assert wb.stats()["num_nodes"] == o.stats()["num_nodes"]

    res = run_ticks_parity(w, [8, w.ticks - 8], dynamic_schema=GRID, box=margin_box(w), graph_ticks=graph_ticks,
                           num_meeting_buckets=1 << 20, extra=extra)
    assert res["meetings"] > 0


def test_grid_rebuild_edges():
    """A point on the high x edge counts in the first cell of the next row
    (the reference's x + 16384*y), as the oracle has it; one on the high y
    edge has no cell and the tick fails with GLINT_ERR_BAD_INPUT."""
    from oracle.oracle import Oracle
    rng = np.random.default_rng(4)
    n, box = 4000, (0.0, 0.0, 64.0, 64.0)
    xy = rng.random((n, 2)) * 63.0
    xy[:40, 0] = 64.0                      # on the high x edge, one cell row above where they lie
    xy[:40, 1] = 10.0 + rng.random(40) * 0.01
    o = Oracle(n, box, 2.0, 8, 5, dynamic_schema=GRID)
    o.tick(xy.reshape(-1), 0)
    cfg = runtime.make_config(n, box, grid_capacity=8, dynamic_schema=GRID, record_contacts=True,
                              num_meeting_buckets=1 << 17, unit_capacity=1 << 20)
    with runtime.Workbench(cfg) as wb:
        wb.tick(xy.reshape(-1), 0)
        b, lv, _ = wb.leaves()
        ob, olv, _ = o.leaves()
        assert np.array_equal(b, ob) and np.array_equal(lv, olv)
        assert np.array_equal(wb.contacts(), o.contacts())
    o.close()
    bad = xy.copy()
    bad[7] = (30.0, 64.0)
    with runtime.Workbench(cfg) as wb:
        with pytest.raises(runtime.GlintError) as e:
            wb.tick(bad.reshape(-1), 0)
        assert e.value.status == runtime.GLINT_ERR_BAD_INPUT and "high y edge" in str(e.value)
