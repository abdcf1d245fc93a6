"""The oracle's FP64 predicates pinned by closed forms and exhaustive checks.

- Point::distance(p, q) <= r (reference geometry.h:40-48, refinement.cpp:33) is
  evaluated in Python's IEEE doubles (no fused multiply-add) as an independent
  computation, on boundary cases at exactly r and one ulp either side, at 30 km
  coordinate magnitudes.
- The device form s <= T with T = glint_reach_threshold(r) is checked to be the
  same predicate as sqrt(s) <= r on every double in a window around T.
- box::distance (geometry.h:174-185) against the clamped-offset closed form.
"""
import math
import random

import numpy as np
import pytest

from glint_pipeline_amd import runtime
from oracle import oracle


def py_within(x1, y1, x2, y2, r):
    dx = x1 - x2
    dy = y1 - y2
    return math.sqrt(dx * dx + dy * dy) <= r


@pytest.mark.parametrize("r", [1.5, 2.0, 3.0, 0.1, 7.25])
def test_point_within_exact_boundary(r):
    # exactly r along an axis is inside; one ulp beyond is outside
    assert oracle.point_within(0.0, 0.0, r, 0.0, r)
    assert not oracle.point_within(0.0, 0.0, math.nextafter(r, math.inf), 0.0, r)
    assert oracle.point_within(0.0, 0.0, math.nextafter(r, 0.0), 0.0, r)
    # 3-4-5 triangle scaled to r: the rounded distance is exactly r
    a, b = 0.6 * r, 0.8 * r
    assert oracle.point_within(10.0, 20.0, 10.0 + a, 20.0 + b, r) == py_within(10.0, 20.0, 10.0 + a, 20.0 + b, r)


def test_point_within_matches_python_at_30km():
    rng = random.Random(7)
    for _ in range(20000):
        x1 = rng.uniform(0, 30000.0)
        y1 = rng.uniform(0, 30000.0)
        ang = rng.uniform(0, 2 * math.pi)
        d = 2.0 + rng.choice([0.0, 1e-13, -1e-13, 1e-9, -1e-9, rng.uniform(-0.5, 0.5)])
        x2, y2 = x1 + d * math.cos(ang), y1 + d * math.sin(ang)
        assert oracle.point_within(x1, y1, x2, y2, 2.0) == py_within(x1, y1, x2, y2, 2.0)


@pytest.mark.parametrize("r", [1.5, 2.0, 3.0, 0.5, 123.456])
def test_reach_threshold_is_the_sqrt_predicate(r):
    t = runtime.lib().glint_reach_threshold(r)
    assert math.sqrt(t) <= r < math.sqrt(math.nextafter(t, math.inf))
    # every double in a window of 20000 ulps around T: s <= T  <=>  sqrt(s) <= r
    s = t
    for _ in range(10000):
        s = math.nextafter(s, 0.0)
    for _ in range(20000):
        assert (s <= t) == (math.sqrt(s) <= r)
        s = math.nextafter(s, math.inf)


def test_reach_threshold_random_squared_distances():
    t = runtime.lib().glint_reach_threshold(2.0)
    rng = np.random.default_rng(3)
    d = rng.uniform(1.9, 2.1, 200000)
    ang = rng.uniform(0, 2 * np.pi, d.size)
    base = rng.uniform(0, 30000, (d.size, 2))
    dx = base[:, 0] - (base[:, 0] + d * np.cos(ang))
    dy = base[:, 1] - (base[:, 1] + d * np.sin(ang))
    s = dx * dx + dy * dy
    assert np.array_equal(s <= t, np.sqrt(s) <= 2.0)


def test_box_distance_closed_form():
    rng = random.Random(11)
    for _ in range(5000):
        lx, ly = rng.uniform(0, 1000), rng.uniform(0, 1000)
        w = rng.choice([3.66, 50.0, 117.1875, 937.5])
        hx, hy = lx + w, ly + w
        px, py = rng.uniform(lx - 10, hx + 10), rng.uniform(ly - 10, hy + 10)
        ddx = max(lx - px, 0.0, px - hx)
        ddy = max(ly - py, 0.0, py - hy)
        assert oracle.box_distance(lx, ly, hx, hy, px, py) == pytest.approx(math.hypot(ddx, ddy), abs=1e-9)
    # inside and on the boundary the distance is exactly 0 (box::contain first)
    assert oracle.box_distance(0, 0, 10, 10, 10, 10) == 0.0
    assert oracle.box_distance(0, 0, 10, 10, 0, 5) == 0.0
    assert oracle.box_distance(0, 0, 10, 10, 12, 10) == 2.0
