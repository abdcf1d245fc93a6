"""The elastic schema update (geometry.cu:1193-1234 collect, 1114-1158 split,
1160-1191 merge) pinned by hand-crafted count sequences on a known tree:
root [0,64]^2 split to 16 leaves of side 16 (3 points each, capacity 4)."""
import numpy as np
import pytest

from oracle.oracle import Oracle


def sixteen_leaf_tree(cap=4, delay=1, r=2.0):
    pts = []
    for i in range(4):
        for j in range(4):
            for k in range(3):
                pts.append((i * 16 + 2 + 4 * k, j * 16 + 3))
    xy = np.array(pts, dtype=np.float64)
    o = Oracle(len(xy), (0, 0, 64, 64), r, cap, schema_update_delay=delay)
    o.build(xy.reshape(-1))
    boxes, levels, _ = o.leaves()
    assert len(boxes) == 16 and set(levels.tolist()) == {2}
    return o


def test_overfull_leaf_splits_one_level():
    o = sixteen_leaf_tree()
    counts = np.full(16, 1, dtype=np.uint32)
    counts[0] = 100                     # far over capacity: still only one level
    o.update_schema_counts(counts)
    boxes, levels, _ = o.leaves()
    assert len(boxes) == 19
    assert boxes[:4].tolist() == [[0, 0, 8, 8], [8, 0, 16, 8], [0, 8, 8, 16], [8, 8, 16, 16]]
    assert levels[:4].tolist() == [3, 3, 3, 3]
    o.close()


def test_capacity_is_strict_for_updates():
    o = sixteen_leaf_tree()
    counts = np.full(16, 4, dtype=np.uint32)   # == cap: update rule needs > cap
    o.update_schema_counts(counts)
    assert len(o.leaves()[0]) == 16
    o.close()


def test_underfull_quadrant_merges():
    o = sixteen_leaf_tree()
    counts = np.full(16, 5, dtype=np.uint32)
    counts[4:8] = [0, 1, 1, 1]          # the second quadrant's four leaves sum to 3 < 4
    # no leaf is over capacity except those at 5 > 4: they split
    o.update_schema_counts(counts)
    boxes, levels, _ = o.leaves()
    assert [32.0, 0.0, 64.0, 32.0] in boxes.tolist()       # merged quadrant (bottom-right)
    assert len(boxes) == 12 * 4 + 1
    o.close()


def test_narrow_leaves_never_split():
    o = sixteen_leaf_tree(r=8.0)        # width 16 is not > 2r = 16
    counts = np.full(16, 50, dtype=np.uint32)
    o.update_schema_counts(counts)
    assert len(o.leaves()[0]) == 16
    o.close()


def test_split_delay_needs_consecutive_overflows():
    o = sixteen_leaf_tree(delay=2)
    over = np.full(16, 1, dtype=np.uint32)
    over[3] = 9
    o.update_schema_counts(over)
    assert len(o.leaves()[0]) == 16     # first overflow: counted, not split
    calm = np.full(16, 1, dtype=np.uint32)
    o.update_schema_counts(calm)        # streak broken
    o.update_schema_counts(over)
    assert len(o.leaves()[0]) == 16
    o.update_schema_counts(over)        # second consecutive overflow
    assert len(o.leaves()[0]) == 19
    o.close()


def test_merge_then_split_back():
    o = sixteen_leaf_tree()
    c = np.full(16, 1, dtype=np.uint32)  # every quadrant sums to 4: not < 4, no merge
    o.update_schema_counts(c)
    assert len(o.leaves()[0]) == 16
    c = np.zeros(16, dtype=np.uint32)
    o.update_schema_counts(c)            # all quadrants merge (one level)
    assert len(o.leaves()[0]) == 4
    o.update_schema_counts(np.zeros(4, dtype=np.uint32))   # then the root
    assert len(o.leaves()[0]) == 1
    o.update_schema_counts(np.array([10], dtype=np.uint32))
    assert len(o.leaves()[0]) == 4
    o.close()
