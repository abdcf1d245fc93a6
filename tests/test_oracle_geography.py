"""The oracle's geography=true predicates (lon/lat degrees), pinned.

- degrees per metre of longitude: an independent evaluation of the
  reference's expression (util.h:57-66) in the same operation order, and
  closed forms (the equator, 60 degrees, the 90-degree clamp);
- the pair test (geometry.h:40-48): offsets planted at known metric
  distances, and a pair whose result depends on which object is p1 (the
  latitude that scales dx): the oracle tests an inside pair with the
  smaller id as p1 (refinement.cpp:30) and a border unit with its object;
- box distances (geometry.h:174-185) against a Python restatement;
- split rules: branches of the initial build have width(true)/2 > r
  (QTree.h:158), leaves of an update do not split unless height(true) > 2r;
- brute force: the contact set of a small lon/lat scene equals every pair
  found by enumerating all leaves under last_valid (with the degree buffers
  of filter.cpp:30-31) and testing with the reference's p1.
"""
import math

import numpy as np
import pytest

from oracle.oracle import (Oracle, box_distance_geo, degree_per_meter_longitude, pair_key,
                           point_within_geo)
from tests.test_oracle_candidates import tree_nodes

PI = 3.14159265
DLAT = 360.0 / (40076.0 * 1000.0)


def dpm_lon(lat: float) -> float:
    a = abs(lat)
    if a == 90:
        a = 89
    return 360.0 / (math.sin((90 - a) * PI / 180) * 40076.0 * 1000.0)


def within_geo(x1, y1, x2, y2, r):
    dx, dy = x1 - x2, y1 - y2
    dy = dy / DLAT
    dx = dx / dpm_lon(y1)
    return math.sqrt(dx * dx + dy * dy) <= r


def test_degrees_per_metre_of_longitude():
    rng = np.random.default_rng(1)
    for lat in list(rng.uniform(-89.9, 89.9, 2000)) + [0.0, 90.0, -90.0, 41.85, -33.9]:
        assert degree_per_meter_longitude(lat) == dpm_lon(lat)
    # the equator: 360 / 40076 km (sin of 90 degrees with PI = 3.14159265 is 1 to 1e-17)
    assert abs(degree_per_meter_longitude(0.0) / (360.0 / 40076000.0) - 1) < 1e-15
    # 60 degrees: twice the equator's (PI's error is 3.6e-9 in the angle)
    assert abs(degree_per_meter_longitude(60.0) / degree_per_meter_longitude(0.0) - 2) < 1e-7
    assert degree_per_meter_longitude(90.0) == degree_per_meter_longitude(89.0)


def test_pair_test_at_known_distances():
    rng = np.random.default_rng(2)
    r = 2.0
    for _ in range(2000):
        x1, y1 = rng.uniform(-120, 120), rng.uniform(-70, 70)
        th = rng.uniform(0, 2 * math.pi)
        d = r + rng.choice([-1e-4, 1e-4]) * rng.uniform(1, 10)
        x2 = x1 + d * math.cos(th) * dpm_lon(y1)
        y2 = y1 + d * math.sin(th) * DLAT
        assert point_within_geo(x1, y1, x2, y2, r) == within_geo(x1, y1, x2, y2, r)
        assert point_within_geo(x1, y1, x2, y2, r) == (d < r)


def asymmetric_pair(r=2.0, lat=50.0):
    """Two points whose distance scaled at the first one's latitude is within
    r and at the second one's is not (bisection on the longitude offset)."""
    y1, y2 = lat, lat - 1.2 * DLAT     # 1.2 m south: a degree of longitude is longer there
    lo, hi = 0.0, 3.0 * dpm_lon(y1)
    for _ in range(200):
        mid = (lo + hi) / 2
        a = within_geo(0.0, y1, mid, y2, r)
        b = within_geo(mid, y2, 0.0, y1, r)
        if a and not b:
            return (0.0, y1), (mid, y2)
        if a:
            lo = mid
        else:
            hi = mid
    raise AssertionError("no asymmetric pair found")


def test_pair_test_uses_p1_latitude():
    p, q = asymmetric_pair()
    assert point_within_geo(*p, *q, 2.0) and not point_within_geo(*q, *p, 2.0)
    box = (-0.001, 49.999, 0.001, 50.001)
    for first, contact in ((p, True), (q, False)):   # pid 0 is p1 of the inside pair
        other = q if first is p else p
        xy = np.array([first, other], dtype=np.float64).reshape(-1)
        o = Oracle(2, box, 2.0, 50, 5, geography=True)
        o.tick(xy, 0)
        assert (len(o.contacts()) == 1) == contact
        o.close()


def test_box_distance():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        lx, ly = rng.uniform(-100, 100), rng.uniform(-60, 60)
        hx, hy = lx + rng.uniform(0, 1e-3), ly + rng.uniform(0, 1e-3)
        px, py = lx + rng.uniform(-2e-3, 3e-3), ly + rng.uniform(-2e-3, 3e-3)
        if lx <= px <= hx and ly <= py <= hy:
            want = 0.0
        else:
            dx = max(abs(px - (lx + hx) / 2) - (hx - lx) / 2, 0.0)
            dy = max(abs(py - (ly + hy) / 2) - (hy - ly) / 2, 0.0)
            dy = dy / DLAT
            dx = dx / dpm_lon(py)
            want = math.sqrt(dx * dx + dy * dy)
        assert box_distance_geo(lx, ly, hx, hy, px, py) == want


def scene(n, seed, lon0=-87.70, lat0=41.85, side_m=300.0):
    """Objects on a 50 m street grid (in metres) near (lon0, lat0), as degrees."""
    rng = np.random.default_rng(seed)
    m = rng.uniform(0, side_m, (n, 2))
    snap = rng.random(n) < 0.5
    m[snap, 0] = np.round(m[snap, 0] / 50) * 50
    m[~snap, 1] = np.round(m[~snap, 1] / 50) * 50
    m += rng.normal(0, 0.6, (n, 2))
    lon = lon0 + m[:, 0] * dpm_lon(lat0)
    lat = lat0 + m[:, 1] * DLAT
    box = (lon0 - 10 * dpm_lon(lat0), lat0 - 10 * DLAT, lon0 + (side_m + 10) * dpm_lon(lat0),
           lat0 + (side_m + 10) * DLAT)
    return np.stack([lon, lat], 1), box


def test_split_rules_in_metres():
    xy, box = scene(3000, 5)
    r, cap = 2.0, 8
    o = Oracle(len(xy), box, r, cap, 5, geography=True)
    o.tick(xy.reshape(-1), 0)
    boxes, levels, counts, nodes = tree_nodes(o, box)
    for b, lvl, k0, k1 in nodes:
        inside = ((xy[:, 0] > b[0]) | (b[0] == box[0])) & (xy[:, 0] <= b[2]) & \
                 ((xy[:, 1] > b[1]) | (b[1] == box[1])) & (xy[:, 1] <= b[3])
        width_m = (b[2] - b[0]) / dpm_lon(b[1])
        if k1 - k0 > 1:   # a branch of the build: count >= cap and width(true) / 2 > r
            assert inside.sum() >= cap and width_m / 2 > r
        else:
            assert inside.sum() < cap or not width_m / 2 > r
    o.close()


def brute_contacts(xy, box, o, r):
    """Every pair the reference tests and finds: inside pairs with the smaller
    id as p1; each object against every object of each leaf that the lookup
    from last_valid accepts (enumerated over all leaves under it), itself p1."""
    boxes, levels, counts, nodes = tree_nodes(o, box)
    leaves = [n for n in nodes if n[3] - n[2] == 1 and tuple(boxes[n[2]]) == n[0]]
    xb, yb = r * dpm_lon(box[1]), r * DLAT

    def bdist(b, x, y):
        lx, ly, hx, hy = b
        if lx <= x <= hx and ly <= y <= hy:
            return 0.0
        dx = max(abs(x - (lx + hx) / 2) - (hx - lx) / 2, 0.0) / dpm_lon(y)
        dy = max(abs(y - (ly + hy) / 2) - (hy - ly) / 2, 0.0) / DLAT
        return math.sqrt(dx * dx + dy * dy)

    def path(x, y):
        out = [[n for n in nodes if n[1] == 0][0]]
        while not (out[-1][3] - out[-1][2] == 1 and tuple(boxes[out[-1][2]]) == out[-1][0]):
            lx, ly, hx, hy = out[-1][0]
            mx, my = (hx + lx) / 2, (hy + ly) / 2
            c = int(y > my) * 2 + int(x > mx)
            cb = [(lx, ly, mx, my), (mx, ly, hx, my), (lx, my, mx, hy), (mx, my, hx, hy)][c]
            out.append([n for n in nodes if n[0] == cb and n[1] == out[-1][1] + 1][0])
        return out

    member = {}
    for i, (x, y) in enumerate(xy):
        member.setdefault(path(x, y)[-1][2], []).append(i)
    found = set()
    for k, ids in member.items():
        for a in ids:
            for b in ids:
                if a < b and within_geo(*xy[a], *xy[b], r):
                    found.add((a, b))
    for i, (x, y) in enumerate(xy):
        p = path(x, y)
        lv = p[0]
        for n in p[1:]:
            if x + xb < n[0][2] and y + yb < n[0][3]:
                lv = n
        if lv is p[-1] or len(nodes) == 1:
            continue
        for leaf in leaves:
            k = leaf[2]
            if k == p[-1][2] or not (lv[2] <= k < lv[3]):
                continue
            # every node from last_valid (exclusive) down to the leaf within r
            anc = [n for n in nodes if n[2] <= k < n[3] and n[1] > lv[1]]
            if any(bdist(n[0], x, y) > r for n in anc):
                continue
            lx, ly, hx, hy = leaf[0]
            if not (y < ly or (y < hy and x < lx)):
                continue
            for q in member.get(k, []):
                if within_geo(x, y, *xy[q], r):
                    found.add((min(i, q), max(i, q)))
    return np.array(sorted(pair_key(a, b) for a, b in found), dtype=np.uint64)


@pytest.mark.parametrize("seed,cap", [(7, 6), (8, 12)])
def test_contacts_equal_brute_force(seed, cap):
    xy, box = scene(900, seed, side_m=150.0)
    o = Oracle(len(xy), box, 2.0, cap, 5, geography=True)
    o.tick(xy.reshape(-1), 0)
    want = brute_contacts(xy, box, o, 2.0)
    assert len(want) > 100 and np.array_equal(o.contacts(), want)
    o.close()


def test_grid_rebuild_refused_with_geography():
    xy, box = scene(50, 9)
    with pytest.raises(ValueError):
        Oracle(len(xy), box, 2.0, 8, 5, geography=True, dynamic_schema=2)
