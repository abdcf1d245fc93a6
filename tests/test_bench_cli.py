"""bench.py's host-side pieces that need no GPU: the committed counter
summaries it reads, and its refusal of more than one GPU."""
import json
import subprocess
import sys
from pathlib import Path

import bench
from glint_pipeline_amd.build import source_digest

ROOT = Path(__file__).resolve().parent.parent


def test_profile_readers_without_a_summary():
    # a workload with no committed summary: no figures, and a note that says so
    vi, slots, note = bench.profiled_valu("k_join", "plain_cpu", 0, 100_000)
    assert vi is None and slots is None and "no counter summary" in note
    t, note = bench.profiled_traffic("k_join", "plain_cpu", 0, 100_000)
    assert t is None and "no counter summary" in note


def test_profile_readers_follow_the_source_digest():
    vi, slots, note = bench.profiled_valu("k_join", "skewed", 1, 10_000_000)
    t, tnote = bench.profiled_traffic("k_join", "skewed", 1, 10_000_000)
    text = (ROOT / "profiles" / "valu_skewed_10m_fp32guard.txt").read_text()
    measured = [ln.split()[1] for ln in text.splitlines() if ln.startswith("SOURCE_SHA")]
    if measured and measured[-1] == source_digest():
        assert vi and vi > 0 and slots and slots > 0
        assert t and t > 0
    else:   # sources changed since the summary: the figures are withheld
        assert vi is None and t is None
        assert "built sources are" in note and "built sources are" in tnote


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode != 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in json.dumps(line)


def test_lonlat_mapping_uses_the_reference_scales():
    from oracle.oracle import degree_per_meter_longitude
    geo = bench.LonLat(-87.63, 41.88)
    assert geo.fx == degree_per_meter_longitude(41.88)
    assert geo.fy == 360.0 / (40076.0 * 1000.0)
    ll = geo.to_degrees([[0.0, 0.0], [1000.0, 2000.0]])
    assert ll[0].tolist() == [-87.63, 41.88]
    assert ll[1].tolist() == [-87.63 + 1000.0 * geo.fx, 41.88 + 2000.0 * geo.fy]

    class W:
        side = 30000.0
    lx, ly, hx, hy = geo.box(W)
    assert lx < -87.63 < hx and ly < 41.88 < hy   # the walk's box [0, side] lies inside
    assert hx > -87.63 + 30000.0 * geo.fx and hy > 41.88 + 30000.0 * geo.fy
