"""The oracle's quadtree pinned against an incremental restatement of the
reference's QTNode::insert/split (QTree.h:157-214) written as a different
procedure (one point at a time with per-leaf lists), and against invariants:
leaves tile the root, every branch holds >= grid_capacity points and is wider
than 2r, every leaf does not, midline points go low/left, and the shape does
not depend on the order of the object ids."""
import numpy as np
import pytest

from oracle.oracle import Oracle


class IncNode:
    """QTree.h QTNode: a leaf splits as soon as it holds >= cap points and
    width/2 > r; children (lo,mid), (mid,hi); child (y>my)*2 + (x>mx)."""

    def __init__(self, lx, ly, hx, hy, level):
        self.b = (lx, ly, hx, hy)
        self.mx, self.my = (hx + lx) / 2, (hy + ly) / 2
        self.level = level
        self.children = None
        self.pts = []

    def insert(self, xy, i, cap, r):
        if self.children is not None:
            x, y = xy[i]
            self.children[int(y > self.my) * 2 + int(x > self.mx)].insert(xy, i, cap, r)
            return
        self.pts.append(i)
        self.split(xy, cap, r)

    def split(self, xy, cap, r):
        lx, ly, hx, hy = self.b
        if not (len(self.pts) >= cap and (hx - lx) / 2 > r):
            return
        mx, my, lv = self.mx, self.my, self.level + 1
        self.children = [IncNode(lx, ly, mx, my, lv), IncNode(mx, ly, hx, my, lv), IncNode(lx, my, mx, hy, lv),
                         IncNode(mx, my, hx, hy, lv)]
        for i in self.pts:
            x, y = xy[i]
            self.children[int(y > my) * 2 + int(x > mx)].pts.append(i)
        self.pts = []
        for c in self.children:
            c.split(xy, cap, r)

    def leaves(self):
        if self.children is None:
            return [(self.b, self.level, len(self.pts))]
        out = []
        for c in self.children:
            out += c.leaves()
        return out


def incremental_tree(xy, box, cap, r):
    root = IncNode(*box, 0)
    for i in range(len(xy)):
        root.insert(xy, i, cap, r)
    return root.leaves()


def oracle_leaves(xy, box, cap, r):
    o = Oracle(len(xy), box, r, cap)
    o.tick(xy.reshape(-1), 0)
    boxes, levels, counts = o.leaves()
    o.close()
    return boxes, levels, counts


def random_points(n, side, seed, clustered=False):
    rng = np.random.default_rng(seed)
    if clustered:
        c = rng.uniform(0, side, (4, 2))
        xy = c[rng.integers(0, 4, n)] + rng.normal(0, side / 40, (n, 2))
    else:
        xy = rng.uniform(0, side, (n, 2))
    xy = np.clip(xy, 0, side)
    # a share of points exactly on street lines / midlines (multiples of side/64)
    k = n // 5
    xy[:k, 0] = np.round(xy[:k, 0] / (side / 64)) * (side / 64)
    return xy


@pytest.mark.parametrize("n,cap,seed,clustered", [(300, 8, 1, False), (800, 16, 2, True), (1500, 32, 3, True),
                                                  (200, 200, 4, False), (1000, 4, 5, True)])
def test_build_matches_incremental_insert(n, cap, seed, clustered):
    side = 1000.0
    xy = random_points(n, side, seed, clustered)
    inc = incremental_tree(xy, (0.0, 0.0, side, side), cap, 2.0)
    boxes, levels, counts = oracle_leaves(xy, (0.0, 0.0, side, side), cap, 2.0)
    assert len(inc) == len(boxes)
    for (b, lv, c), ob, olv, oc in zip(inc, boxes, levels, counts):
        assert tuple(ob) == b and olv == lv and oc == c


def test_leaves_tile_the_root_and_respect_the_split_rule():
    side, cap, r = 2000.0, 12, 2.0
    xy = random_points(3000, side, 9, True)
    boxes, levels, counts = oracle_leaves(xy, (0.0, 0.0, side, side), cap, r)
    area = ((boxes[:, 2] - boxes[:, 0]) * (boxes[:, 3] - boxes[:, 1])).sum()
    assert area == pytest.approx(side * side, rel=1e-12)
    assert counts.sum() == len(xy)
    # a leaf never satisfies the split rule
    w = boxes[:, 2] - boxes[:, 0]
    assert not np.any((counts >= cap) & (w / 2 > r))
    # parents of leaves (4 siblings, same level, same parent box) satisfy it
    for i in range(len(boxes) - 3):
        if levels[i] > 0 and all(levels[i:i + 4] == levels[i]):
            b = boxes[i:i + 4]
            if b[0, 2] == b[1, 0] and b[0, 3] == b[2, 1] and b[3, 0] == b[0, 2] and b[3, 1] == b[0, 3]:
                pw = b[1, 2] - b[0, 0]
                assert counts[i:i + 4].sum() >= cap and pw / 2 > r


def test_midline_points_go_low_left():
    side = 400.0
    xy = np.array([[200.0, 200.0], [200.0, 100.0], [100.0, 200.0], [300.0, 300.0], [50.0, 50.0],
                   [250.0, 50.0], [50.0, 250.0], [350.0, 350.0]])
    o = Oracle(len(xy), (0, 0, side, side), 2.0, 2)
    o.tick(xy.reshape(-1), 0)
    boxes, _, _ = o.leaves()
    for x, y in xy:
        leaf, _, _ = o.locate(x, y)
        lx, ly, hx, hy = boxes[leaf]
        # inside the box, and on a shared edge the point belongs to the lower/left box
        assert lx <= x <= hx and ly <= y <= hy
        if x == lx and lx > 0:
            raise AssertionError("a point on a vertical midline went right")
        if y == ly and ly > 0:
            raise AssertionError("a point on a horizontal midline went up")
    o.close()


def test_shape_independent_of_object_order():
    side = 1500.0
    xy = random_points(2000, side, 13, True)
    perm = np.random.default_rng(1).permutation(len(xy))
    a = oracle_leaves(xy, (0.0, 0.0, side, side), 10, 2.0)
    b = oracle_leaves(xy[perm], (0.0, 0.0, side, side), 10, 2.0)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
