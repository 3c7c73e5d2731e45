"""Host tree invariants, checked by brute force (the tree replaces the absent simpletree.BalancedTree)."""

import numpy as np
import pytest

from conftest import annulus_points, sphere_points
from skelfmm_amd.tree import BalancedTree, findin_sorted, morton_decode, morton_encode


def _trees():
    rng = np.random.default_rng(0)
    yield BalancedTree(rng.random((3000, 2)), leaf_size=40)
    yield BalancedTree(annulus_points(4000), leaf_size=50)
    yield BalancedTree(rng.random((4000, 3)), leaf_size=60)
    yield BalancedTree(sphere_points(5000), leaf_size=64)
    # strongly clustered cloud: forces adaptivity and balancing
    c = np.vstack([rng.random((2000, 3)) * 0.01 + 0.2, rng.random((300, 3))])
    yield BalancedTree(c, leaf_size=50)


TREES = list(_trees())


def _bounds(t, b):
    c = t.get_box_center(b)
    h = 0.5 * t.get_box_length(b)
    return c - h, c + h


def test_morton_roundtrip():
    rng = np.random.default_rng(1)
    for d, nb in ((2, 30), (3, 20)):
        ic = rng.integers(0, 2 ** nb, size=(500, d))
        assert np.array_equal(morton_decode(morton_encode(ic, nb), nb, d), ic)
    assert findin_sorted(np.array([1, 4, 9]), 4) == 1
    assert findin_sorted(np.array([1, 4, 9]), 5) == -1


@pytest.mark.parametrize("t", TREES)
def test_leaves_partition_points_and_respect_leaf_size(t):
    seen = np.zeros(t.N, dtype=int)
    for leaf in t.get_leaves():
        inds = t.get_box_inds(leaf)
        assert 0 < inds.shape[0] <= t.leaf_size
        seen[inds] += 1
        lo, hi = _bounds(t, leaf)
        P = t.XX[inds]
        assert np.all(P >= lo - 1e-12) and np.all(P <= hi + 1e-12)
    assert np.all(seen == 1)


@pytest.mark.parametrize("t", TREES)
def test_numbering_levels_and_children(t):
    assert t.get_box_level(0) == 0
    for lev in range(t.nlevels):
        for b in t.get_boxes_level(lev):
            assert t.get_box_level(b) == lev
    for b in range(t.nboxes):
        ch = t.get_box_children(b)
        if t.is_leaf(b):
            assert ch.size == 0
        else:
            # children hold exactly the parent's points
            pts = np.sort(np.concatenate([t.get_box_inds(c) for c in ch]))
            assert np.array_equal(pts, np.sort(t.get_box_inds(b)))
            for c in ch:
                assert t.get_box_parent(c) == b and t.get_box_level(c) == t.get_box_level(b) + 1
    assert t.nleaves == t.get_leaves().shape[0]
    assert t.level_sep[-1] == t.nboxes


def _touch(t, a, b):
    la, ha = _bounds(t, a)
    lb, hb = _bounds(t, b)
    tol = 1e-9 * t.L0
    return np.all(la <= hb + tol) and np.all(lb <= ha + tol)


@pytest.mark.parametrize("t", TREES)
def test_colleagues_brute_force(t):
    for lev in range(t.nlevels):
        boxes = t.get_boxes_level(lev)
        rows = t.get_boxes_colleague_neigh_fullvec(boxes)
        for b, row in zip(boxes, rows):
            got = set(row[row >= 0].tolist())
            want = {int(c) for c in boxes if _touch(t, b, c)}
            assert got == want
            assert int(b) in got


@pytest.mark.parametrize("t", TREES)
def test_coarse_neighbors_brute_force(t):
    leaves = t.get_leaves()
    for lev in range(1, t.nlevels):
        boxes = t.get_boxes_level(lev)
        rows = t.get_boxes_coarse_neigh_fullvec(boxes)
        for b, row in zip(boxes, rows):
            got = set(row[row >= 0].tolist())
            want = {int(c) for c in leaves if t.get_box_level(c) < lev and _touch(t, b, c)}
            assert got == want


@pytest.mark.parametrize("t", TREES)
def test_two_to_one_balance(t):
    leaves = t.get_leaves()
    lv = np.array([t.get_box_level(b) for b in leaves])
    for i, a in enumerate(leaves):
        for j in range(i + 1, len(leaves)):
            if abs(lv[i] - lv[j]) > 1 and _touch(t, a, leaves[j]):
                raise AssertionError(f"leaves {a} and {leaves[j]} touch across {abs(lv[i] - lv[j])} levels")


def test_leaves_above_and_bad_inputs():
    t = TREES[4]
    for lev in range(t.nlevels + 1):
        la = t.get_leaves_above(lev)
        assert all(t.get_box_level(b) < lev for b in la)
    with pytest.raises(ValueError):
        BalancedTree(np.zeros((0, 3)), leaf_size=4)
    with pytest.raises(ValueError):
        BalancedTree(np.zeros((10, 4)), leaf_size=4)
    # more coincident points than a leaf can hold cannot be split
    with pytest.raises(ValueError):
        BalancedTree(np.zeros((10, 2)), leaf_size=4)
    one = BalancedTree(np.array([[0.3, 0.4, 0.5]]), leaf_size=1)
    assert one.nboxes == 1 and one.nleaves == 1
