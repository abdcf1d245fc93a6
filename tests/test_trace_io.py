"""Reading the reference's trace files and the lon/lat -> metres conversion
(host side; no GPU).

- trace files (u32 objects, u32 ticks, box, tick-major Points;
  process.cpp:146-187): a known-answer file whose every point encodes its
  tick and object, read with fewer, equal and more objects than it holds
  (more: the next objects come from the file's later ticks);
- chunk files "<prefix><st>_<N>.tr" (datagen.cpp:18-26);
- errors for missing, short and out-of-range reads (the reference asserts);
- the conversion factors against util.h's formulas evaluated in Python,
  including the 90-degree clamp."""
import math

import numpy as np
import pytest

from glint_pipeline_amd import runtime

STATUS = runtime.STATUS_NAMES


def kat_points(nfile, duration):
    t = np.arange(duration, dtype=np.float64)[:, None]
    o = np.arange(nfile, dtype=np.float64)[None, :]
    return np.stack([t * 100 + o, -t - o / 1000], axis=-1)   # (duration, nfile, 2)


def write_trace(path, pts, box=(1.0, 2.0, 3.0, 4.0)):
    d, n, _ = pts.shape
    with open(path, "wb") as f:
        f.write(np.array([n, d], dtype=np.uint32).tobytes())
        f.write(np.array(box, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(pts, dtype=np.float64).tobytes())


def test_header_and_plain_window(tmp_path):
    pts = kat_points(4, 12)
    p = tmp_path / "t.tr"
    write_trace(p, pts, box=(-87.9, 41.6, -87.5, 42.0))
    h = runtime.tr_header(p)
    assert (h["num_objects"], h["duration"]) == (4, 12)
    assert (h["low_x"], h["low_y"], h["high_x"], h["high_y"]) == (-87.9, 41.6, -87.5, 42.0)
    got = runtime.tr_read(p, 4, 3, 5)
    assert np.array_equal(got, pts[3:8])
    got = runtime.tr_read(p, 2, 0, 12)        # fewer objects: the first ones of each tick
    assert np.array_equal(got, pts[:, :2])


def test_more_objects_than_the_file_holds(tmp_path):
    # 3 objects in the file, 7 asked for ticks 1..2: objects 0-2 from ticks
    # 1-2, objects 3-5 from ticks 3-4, object 6 from ticks 5-6 (its object 0)
    pts = kat_points(3, 10)
    p = tmp_path / "t.tr"
    write_trace(p, pts)
    got = runtime.tr_read(p, 7, 1, 2)
    assert got.shape == (2, 7, 2)
    want_tick = [[1, 1, 1, 3, 3, 3, 5], [2, 2, 2, 4, 4, 4, 6]]
    want_obj = [0, 1, 2, 0, 1, 2, 0]
    for i in range(2):
        for j in range(7):
            assert got[i, j, 0] == want_tick[i][j] * 100 + want_obj[j], (i, j)
            assert got[i, j, 1] == -want_tick[i][j] - want_obj[j] / 1000


def test_errors(tmp_path):
    pts = kat_points(3, 10)
    p = tmp_path / "t.tr"
    write_trace(p, pts)
    with pytest.raises(runtime.GlintError) as e:
        runtime.tr_read(tmp_path / "missing.tr", 3, 0, 1)
    assert STATUS[e.value.status] == "GLINT_ERR_BAD_INPUT"
    with pytest.raises(runtime.GlintError) as e:
        runtime.tr_read(p, 3, 8, 3)           # ticks past the file's duration
    assert STATUS[e.value.status] == "GLINT_ERR_INVALID_ARGUMENT"
    with pytest.raises(runtime.GlintError) as e:
        runtime.tr_read(p, 7, 0, 5)           # 7 objects x 5 ticks > 3 x 10
    assert STATUS[e.value.status] == "GLINT_ERR_INVALID_ARGUMENT"
    short = tmp_path / "short.tr"
    short.write_bytes(p.read_bytes()[:-16])
    with pytest.raises(runtime.GlintError) as e:
        runtime.tr_read(short, 3, 9, 1)
    assert STATUS[e.value.status] == "GLINT_ERR_BAD_INPUT"


def test_chunk_files(tmp_path):
    pts = kat_points(5, 100)
    (tmp_path / "run200_5.tr").write_bytes(pts.tobytes())
    got = runtime.tr_read_chunk(str(tmp_path / "run"), 200, 5)
    assert np.array_equal(got, pts)
    with pytest.raises(runtime.GlintError):
        runtime.tr_read_chunk(str(tmp_path / "run"), 300, 5)
    (tmp_path / "run300_5.tr").write_bytes(pts.tobytes()[:-8])
    with pytest.raises(runtime.GlintError):
        runtime.tr_read_chunk(str(tmp_path / "run"), 300, 5)


@pytest.mark.parametrize("lat", [41.6, -33.9, 0.0, 89.5, 90.0, -90.0])
def test_projection_factors(lat):
    p = runtime.projection_for_box(-87.9, lat)
    absla = abs(lat)
    if absla == 90:
        absla = 89
    assert p.deg_per_m_lon == 360.0 / (math.sin((90 - absla) * 3.14159265 / 180) * 40076.0 * 1000.0)
    assert p.deg_per_m_lat == 360.0 / (40076.0 * 1000.0)
    assert (p.lon0, p.lat0) == (-87.9, lat)


def test_projection_rejects_bad_latitudes():
    for lat in (91.0, float("nan")):
        with pytest.raises(runtime.GlintError):
            runtime.projection_for_box(0.0, lat)
