"""Reference lon/lat traces on the GPU: the device conversion to metres is
bit-equal to the same IEEE expressions in numpy, and a reference-format
trace file read, converted on the device and traced gives the oracle's
meetings on the converted coordinates."""
import numpy as np
import pytest

from glint_pipeline_amd import gen, runtime
from tests.parity import run_ticks_parity

pytestmark = pytest.mark.gpu

LON0, LAT0 = -87.94, 41.64    # a box corner near Chicago


def test_device_projection_is_bit_exact():
    rng = np.random.default_rng(7)
    n = 1_000_003
    ll = np.stack([LON0 + rng.random(n) * 0.5, LAT0 + rng.random(n) * 0.4], axis=1)
    p = runtime.projection_for_box(LON0, LAT0)
    want = np.stack([(ll[:, 0] - p.lon0) / p.deg_per_m_lon, (ll[:, 1] - p.lat0) / p.deg_per_m_lat], axis=1)
    buf = runtime.DeviceBuffer(ll.nbytes)
    buf.upload(ll)
    runtime.project_device(p, buf.address(), buf.address(), n)   # in place
    got = buf.download(np.zeros_like(ll))
    buf.free()
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_reference_trace_file_end_to_end(tmp_path):
    w = gen.uniform_city(6000, 30, 600.0, seed=71, grid_capacity=32)
    g = gen.Generator(w)
    # positions() and step() return the generator's own buffer: copy each tick
    xy = np.stack([g.positions().copy()] + [g.step().copy() for _ in range(w.ticks - 1)])
    xy = xy.reshape(w.ticks, w.num_objects, 2)
    p = runtime.projection_for_box(LON0, LAT0)
    lonlat = np.empty_like(xy)
    lonlat[..., 0] = LON0 + xy[..., 0] * p.deg_per_m_lon
    lonlat[..., 1] = LAT0 + xy[..., 1] * p.deg_per_m_lat
    path = tmp_path / "city.tr"
    with open(path, "wb") as f:
        f.write(np.array([w.num_objects, w.ticks], dtype=np.uint32).tobytes())
        f.write(np.array([lonlat[..., 0].min(), lonlat[..., 1].min(), lonlat[..., 0].max(), lonlat[..., 1].max()],
                         dtype=np.float64).tobytes())
        f.write(lonlat.tobytes())
    tr = runtime.tr_read(path, w.num_objects, 0, w.ticks)
    assert np.array_equal(tr, lonlat)
    metres = np.stack([(tr[..., 0] - p.lon0) / p.deg_per_m_lon, (tr[..., 1] - p.lat0) / p.deg_per_m_lat], axis=-1)
    box = (-1.0, -1.0, w.side + 1.0, w.side + 1.0)
    buf = runtime.DeviceBuffer(tr.nbytes)
    buf.upload(tr)
    runtime.project_device(p, buf.address(), buf.address(), w.num_objects * w.ticks)
    # every tick's meetings, the active pairs and the leaves of the projected
    # file equal the oracle's on the host projection of the same file
    res = run_ticks_parity(w, [w.ticks], box=box, precision=runtime.PRECISION_FP64,
                           trace=(buf, lambda t: np.ascontiguousarray(metres[t]).reshape(-1)))
    buf.free()
    assert res["meetings"] > 20
