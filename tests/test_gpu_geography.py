"""Geography mode on the GPU (the reference's geography=true): positions are
longitude/latitude degrees, the reach and the split rules are metres.
Every tick's meetings, contacts, tested pairs and leaves equal the oracle's
in the same mode:

- the device's degrees per metre of longitude are the reference's
  expression with the sine correctly rounded, which is the host libm's
  result wherever the libm's sine is correctly rounded;
- a reference-format trace file of lon/lat degrees read back and run through
  glint_run_ticks on the raw degrees (graph replays, meeting log);
- tick by tick with recorded contacts at three latitudes (north, south, far
  north), with dense hotspots whose leaves cannot split below 2 r and so
  span several join tiles;
- the configurations the reference cannot run in this mode are refused.
"""
import math
from decimal import Decimal, localcontext

import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime
from oracle.oracle import degree_per_meter_longitude
from tests.parity import run_parity, run_ticks_parity

pytestmark = pytest.mark.gpu

DLAT = 360.0 / (40076.0 * 1000.0)


def lonlat_walk(w: gen.Workload, ticks: int, lon0: float, lat0: float):
    """The workload's walk in metres mapped to degrees at (lon0, lat0); the
    box is the walk's box with 1 m of margin, in degrees."""
    g = gen.Generator(w)
    xy = np.stack([g.positions().copy()] + [g.step().copy() for _ in range(ticks - 1)])
    xy = xy.reshape(ticks, w.num_objects, 2)
    f = degree_per_meter_longitude(lat0)
    ll = np.empty_like(xy)
    ll[..., 0] = lon0 + xy[..., 0] * f
    ll[..., 1] = lat0 + xy[..., 1] * DLAT
    box = (lon0 - f, lat0 - DLAT, lon0 + (w.side + 1) * f, lat0 + (w.side + 1) * DLAT)
    return ll, box


def sin_correctly_rounded(x: float) -> float:
    """sin(x) rounded once to the nearest double: the Taylor series in
    60-digit decimal arithmetic from the exact value of x."""
    with localcontext() as ctx:
        ctx.prec = 60
        d = Decimal(x)
        t, s, x2, n = d, d, d * d, 1
        while abs(t) > Decimal(10) ** -55:
            t = -t * x2 / ((n + 1) * (n + 2))
            n += 2
            s += t
        return float(s)


def test_longitude_scale_is_correctly_rounded():
    """The device factor is the reference's expression with its sine
    correctly rounded; it equals the host libm's (the reference's) at every
    latitude where the libm sine is correctly rounded, and the libm is
    correctly rounded nearly everywhere."""
    rng = np.random.default_rng(5)
    lat = np.concatenate([rng.uniform(-89.999, 89.999, 40_000), rng.uniform(40, 42, 10_000),
                          rng.uniform(89, 90, 2000),
                          np.array([0.0, -0.0, 90.0, -90.0, 89.0, -89.0, 45.0, 60.0, 1e-300, 89.99999999])])
    buf, out = runtime.DeviceBuffer(lat.nbytes), runtime.DeviceBuffer(lat.nbytes)
    buf.upload(lat)
    runtime.longitude_scale_device(buf.address(), out.address(), len(lat))
    got = out.download(np.zeros_like(lat))
    buf.free()
    out.free()
    a = np.abs(lat)
    a[a == 90] = 89
    x = (90 - a) * 3.14159265 / 180
    sin_cr = np.array([sin_correctly_rounded(float(v)) for v in x])
    sin_libm = np.array([math.sin(float(v)) for v in x])
    want = 360.0 / (sin_cr * 40076.0 * 1000.0)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    libm = np.array([degree_per_meter_longitude(float(v)) for v in lat])
    same = sin_libm == sin_cr
    assert np.array_equal(got[same].view(np.uint64), libm[same].view(np.uint64))
    assert (~same).sum() < 0.005 * len(lat)


def test_trace_file_of_degrees_through_run_ticks(tmp_path):
    w = gen.uniform_city(6000, 30, 600.0, seed=171, grid_capacity=32)
    ll, box = lonlat_walk(w, w.ticks, -87.94, 41.64)
    path = tmp_path / "city.tr"
    with open(path, "wb") as f:
        f.write(np.array([w.num_objects, w.ticks], dtype=np.uint32).tobytes())
        f.write(np.array(box, dtype=np.float64).tobytes())
        f.write(ll.tobytes())
    hdr = runtime.tr_header(path)
    tr = runtime.tr_read(path, w.num_objects, 0, w.ticks)
    assert np.array_equal(tr, ll)
    box = (hdr["low_x"], hdr["low_y"], hdr["high_x"], hdr["high_y"])
    buf = runtime.DeviceBuffer(tr.nbytes)
    buf.upload(tr)
    res = run_ticks_parity(w, [1, 8, 16, 5], box=box, precision=runtime.PRECISION_FP64, geography=True,
                           graph_ticks=8, trace=(buf, lambda t: np.ascontiguousarray(tr[t]).reshape(-1)))
    buf.free()
    assert res["meetings"] > 20


@pytest.mark.parametrize("lon0,lat0", [(-87.94, 41.64), (151.10, -33.95), (18.90, 69.60)])
def test_hotspots_tick_by_tick_with_contacts(lon0, lat0):
    # 1000 m box: leaves stop splitting at 3.9 m, and the hotspots hold
    # about 2000 objects in such a leaf
    w = gen.Workload("hot_geo", 30_000, 14, 1000.0, seed=172, num_hotspots=3, hotspot_fraction=0.6,
                     hotspot_sigma=5.0, grid_capacity=64)
    ll, box = lonlat_walk(w, w.ticks, lon0, lat0)
    stats = []

    def positions(t):
        return np.ascontiguousarray(ll[t]).reshape(-1)

    tot = run_parity(w, w.ticks, box=box, positions=positions, geography=True, check_leaves_every=3,
                     contact_capacity=1 << 24, meeting_capacity=1 << 22)
    assert tot["contacts"] > 10_000_000 and tot["meetings"] > 100_000
    # leaves narrower than 4 m do not split: some hold more than a join tile
    with runtime.Workbench(runtime.make_config(w.num_objects, box, grid_capacity=w.grid_capacity,
                                               geography=True, precision=runtime.PRECISION_FP64,
                                               contact_capacity=1 << 24)) as wb:
        wb.tick(positions(0), 0)
        stats.append(wb.stats())
    assert stats[0]["max_leaf_count"] > 512


def test_unsupported_configurations_refused():
    box = (-87.94, 41.64, -87.93, 41.65)
    for kw in (dict(precision=runtime.PRECISION_FP32_GUARD),
               dict(precision=runtime.PRECISION_FP64, dynamic_schema=runtime.SCHEMA_GRID_REBUILD),
               dict(precision=runtime.PRECISION_FP64, box=(-87.94, 41.64, -87.93, 90.0))):
        b = kw.pop("box", box)
        with pytest.raises(runtime.GlintError) as e:
            runtime.Workbench(runtime.make_config(100, b, geography=True, **kw))
        assert e.value.status == runtime.GLINT_ERR_INVALID_ARGUMENT


def test_lonlat_cities_in_one_batch():
    """Three lon/lat cities (north, south, far north) in one batch share the
    geography kernels' fused launches; each equals itself run alone."""
    T = 16
    places = [(-87.94, 41.64), (151.10, -33.95), (18.90, 69.60)]
    ws = [gen.uniform_city(n, T, 400.0, seed=180 + i, grid_capacity=24) for i, n in enumerate((3000, 5000, 8000))]
    cfgs, bufs, boxes = [], [], []
    for w, (lon0, lat0) in zip(ws, places):
        ll, box = lonlat_walk(w, T, lon0, lat0)
        b = runtime.DeviceBuffer(ll.nbytes)
        b.upload(ll)
        bufs.append(b)
        cfgs.append(runtime.make_config(w.num_objects, box, grid_capacity=w.grid_capacity, geography=True,
                                        precision=runtime.PRECISION_FP64, meeting_log_capacity=1 << 18))
    with runtime.Batch(cfgs, graph_ticks=4) as bt:
        assert bt.groups() == 1
        bt.run_ticks([b.address(0) for b in bufs], 0, T)
        got = [(c.meeting_log(), c.active_pairs(), c.leaves()) for c in bt.contexts]
    total = 0
    for cfg, b, (log, act, lvs) in zip(cfgs, bufs, got):
        with runtime.Workbench(cfg) as wb:
            wb.run_ticks(b.address(0), 0, T)
            assert np.array_equal(log, wb.meeting_log())
            for x, y in zip(act, wb.active_pairs()):
                assert np.array_equal(x, y)
            for x, y in zip(lvs, wb.leaves()):
                assert np.array_equal(x, y)
        total += len(log)
    assert total > 1000
    for b in bufs:
        b.free()
