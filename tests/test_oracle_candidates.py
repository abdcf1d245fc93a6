"""The reference's candidate rule and contact set, pinned.

- The Q1 fixture: a pair 0.89 m apart on opposite sides of a coarse quadrant
  corner that the reference never tests (the top/right rule of
  filter.cpp:139-141 with last_valid from filter.cpp:235-238), while the
  complete mode finds it.
- Brute force: for small trees, the candidate leaves of every point computed by
  enumerating ALL leaves and checking the path condition leaf by leaf, and the
  contact set from all O(N^2) pairs filtered by those candidate sets, against
  the oracle's depth-first lookup and join.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, pair_key


def tree_nodes(o, root):
    boxes, levels, counts = o.leaves()
    nodes = []

    def rec(box, level, k):
        if levels[k] == level and tuple(boxes[k]) == box:
            nodes.append((box, level, k, k + 1))
            return k + 1
        lx, ly, hx, hy = box
        mx, my = (hx + lx) / 2, (hy + ly) / 2
        start = k
        for cb in ((lx, ly, mx, my), (mx, ly, hx, my), (lx, my, mx, hy), (mx, my, hx, hy)):
            k = rec(cb, level + 1, k)
        nodes.append((box, level, start, k))
        return k

    rec(tuple(root), 0, 0)
    return boxes, levels, counts, nodes


def box_dist(b, x, y):
    lx, ly, hx, hy = b
    if lx <= x <= hx and ly <= y <= hy:
        return 0.0
    dx = max(abs(x - (lx + hx) / 2) - (hx - lx) / 2, 0.0)
    dy = max(abs(y - (ly + hy) / 2) - (hy - ly) / 2, 0.0)
    return math.sqrt(dx * dx + dy * dy)


def brute_candidates(x, y, boxes, nodes, r, complete):
    """Every leaf L (enumerated, not searched) that the lookup must return:
    L lies under last_valid, every node from last_valid (exclusive) down to L
    is within r of the point, L is not the point's own leaf, and (reference
    mode) L passes the top/right rule."""
    # descent path and last_valid
    path = []
    cur = [n for n in nodes if n[1] == 0][0]
    while True:
        path.append(cur)
        if cur[3] - cur[2] == 1 and tuple(boxes[cur[2]]) == cur[0]:
            break
        lx, ly, hx, hy = cur[0]
        mx, my = (hx + lx) / 2, (hy + ly) / 2
        c = int(y > my) * 2 + int(x > mx)
        cb = [(lx, ly, mx, my), (mx, ly, hx, my), (lx, my, mx, hy), (mx, my, hx, hy)][c]
        cur = [n for n in nodes if n[0] == cb and n[1] == cur[1] + 1][0]
    own = path[-1][2]
    lv = path[0]
    for n in path[1:]:
        if x + r < n[0][2] and y + r < n[0][3]:
            lv = n
    if complete:
        lv = path[0]
    elif lv is path[-1]:
        return own, set()
    if len(nodes) == 1:
        return own, set()
    out = set()
    for k in range(lv[2], lv[3]):
        if k == own:
            continue
        b = tuple(boxes[k])
        # all nodes strictly below lv on the way down to leaf k
        chain = [n for n in nodes if n[2] <= k < n[3] and n[1] > lv[1]]
        if not all(box_dist(n[0], x, y) <= r for n in chain):
            continue
        lx, ly, hx, hy = b
        if complete or y < ly or (y < hy and x < lx):
            out.add(k)
    return own, out


def q1_points():
    # root [0,40]^2; quadrants [0,20]^2 and [20,40]x[0,20] split (4 points
    # each, one per child), the top quadrants stay leaves (capacity 4)
    return np.array([[20.2, 9.5],    # p, in A = [20,30]x[0,10] (bottom-left child of R)
                     [19.8, 10.3],   # q, in Q = [10,20]x[10,20] (top-right child of [0,20]^2)
                     [5.0, 5.0], [15.0, 5.0], [5.0, 15.0],
                     [35.0, 5.0], [25.0, 15.0], [35.0, 15.0]])


def test_q1_pair_is_missed_by_the_reference_rule():
    xy = q1_points()
    assert math.hypot(*(xy[0] - xy[1])) < 0.9
    o = Oracle(len(xy), (0, 0, 40, 40), 2.0, 4)
    o.tick(xy.reshape(-1), 0)
    boxes, _, _ = o.leaves()
    # the tree has the shape of the worked example
    assert [20.0, 0.0, 30.0, 10.0] in boxes.tolist() and [10.0, 10.0, 20.0, 20.0] in boxes.tolist()
    assert pair_key(0, 1) not in set(o.contacts().tolist())
    # from p: last_valid is R = [20,40]x[0,20] (leaves 4..8 under it), Q not visited
    leaf_p, span_p, cand_p = o.locate(*xy[0])
    assert tuple(boxes[leaf_p]) == (20.0, 0.0, 30.0, 10.0)
    assert all(span_p[0] <= c < span_p[1] for c in cand_p)
    leaf_q, span_q, cand_q = o.locate(*xy[1])
    assert tuple(boxes[leaf_q]) == (10.0, 10.0, 20.0, 20.0)
    assert leaf_p not in cand_q           # A is visited from q but rejected by the rule
    o.close()
    oc = Oracle(len(xy), (0, 0, 40, 40), 2.0, 4, complete=True)
    oc.tick(xy.reshape(-1), 0)
    assert pair_key(0, 1) in set(oc.contacts().tolist())
    oc.close()


@pytest.mark.parametrize("seed,n,cap,complete", [(1, 200, 6, False), (2, 400, 10, False), (3, 300, 5, True),
                                                 (4, 500, 16, False), (5, 250, 4, True)])
def test_candidates_and_contacts_brute_force(seed, n, cap, complete):
    rng = np.random.default_rng(seed)
    side = 64.0
    xy = rng.uniform(0, side, (n, 2))
    xy[: n // 4] = np.round(xy[: n // 4] / 4.0) * 4.0    # points on midlines and corners
    # points exactly r below a leaf edge: x + r and y + r land on box edges
    q = n // 4
    xy[q: 2 * q] = np.round(xy[q: 2 * q] / 2.0) * 2.0
    r = 2.0
    o = Oracle(n, (0, 0, side, side), r, cap, complete=complete)
    o.tick(xy.reshape(-1), 0)
    boxes, levels, counts, nodes = tree_nodes(o, (0.0, 0.0, side, side))
    leaf_of = {}
    cand = {}
    for i, (x, y) in enumerate(xy):
        own, exp = brute_candidates(x, y, boxes, nodes, r, complete)
        leaf, _, got = o.locate(x, y)
        assert leaf == own
        assert set(got) == exp, f"point {i}: {sorted(got)} != {sorted(exp)}"
        leaf_of[i] = own
        cand[i] = exp
    # contacts: all pairs within r that share a leaf or whose leaves are candidates
    expected = set()
    for i in range(n):
        for j in range(i + 1, n):
            dx, dy = xy[i, 0] - xy[j, 0], xy[i, 1] - xy[j, 1]
            if math.sqrt(dx * dx + dy * dy) > r:
                continue
            if leaf_of[i] == leaf_of[j] or leaf_of[j] in cand[i] or leaf_of[i] in cand[j]:
                expected.add(pair_key(i, j))
    got = set(o.contacts().tolist())
    assert got == expected
    if complete:
        allpairs = {pair_key(i, j) for i in range(n) for j in range(i + 1, n)
                    if math.sqrt((xy[i, 0] - xy[j, 0]) ** 2 + (xy[i, 1] - xy[j, 1]) ** 2) <= r}
        assert got == allpairs
    o.close()
