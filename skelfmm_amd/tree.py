"""Adaptive, level-restricted quad/octree over a point cloud.

The reference imports ``BalancedTree`` from the external ``simpletree`` package
(``skelfmm/skel_fmm.py:8``, ``skelfmm/recursive_skel.py:7``), which is not vendored.
This module supplies the same contract as it is used by the reference
(``recursive_skel.py:54-68,142-166,209-262,622-780``):

* ``N``, ``ndim``, ``nlevels``, ``nboxes``, ``nleaves``, ``leaf_size``, ``L0``,
  ``c0``, ``level_sep`` and ``XX`` (points in the caller's order);
* box 0 is the root; boxes are numbered level by level, and inside a level in
  Morton order of their integer coordinates;
* ``get_leaves``, ``get_boxes_level``, ``get_box_inds``, ``is_leaf``,
  ``are_leaves``, ``get_box_children``, ``get_boxes_children_fullvec``,
  ``get_boxes_colleague_neigh_fullvec`` (same level, touching, self included),
  ``get_boxes_coarse_neigh_fullvec`` (touching leaves on coarser levels),
  ``get_leaves_above``, ``get_box_length``, ``get_boxes_centers``,
  ``get_box_center``, ``get_box_level``.

Construction: points are sorted once by a Morton key at depth ``D`` (30 bits
per axis in 2D, 20 in 3D), so every box owns a contiguous range of the sorted
order. A box is split while it holds more than ``leaf_size`` points; empty
children are not created (a leaf never holds zero points, matching the
``assert`` at ``recursive_skel.py:151``). With ``balance=True`` (default) the
tree is then 2:1 level-restricted: any two touching leaves differ by at most
one level.
"""

from __future__ import annotations

import numpy as np

__all__ = ["BalancedTree", "findin_sorted", "morton_encode", "morton_decode"]


def findin_sorted(sorted_arr, value):
    """Index of ``value`` in the sorted array ``sorted_arr`` (``-1`` if absent)."""
    sorted_arr = np.asarray(sorted_arr)
    i = int(np.searchsorted(sorted_arr, value))
    if i < sorted_arr.shape[0] and sorted_arr[i] == value:
        return i
    return -1


def morton_encode(icoords, nbits):
    """Interleave integer coordinates ``(n, d)`` into Morton codes (bit b of axis k -> bit b*d+k)."""
    icoords = np.asarray(icoords, dtype=np.int64)
    n, d = icoords.shape
    code = np.zeros(n, dtype=np.int64)
    for b in range(nbits):
        for k in range(d):
            code |= ((icoords[:, k] >> b) & 1) << (b * d + k)
    return code


def morton_decode(codes, nbits, d):
    codes = np.asarray(codes, dtype=np.int64)
    out = np.zeros((codes.shape[0], d), dtype=np.int64)
    for b in range(nbits):
        for k in range(d):
            out[:, k] |= ((codes >> (b * d + k)) & 1) << b
    return out


class BalancedTree:
    """Adaptive 2^d-tree with at most ``leaf_size`` points per leaf."""

    def __init__(self, XX, leaf_size=200, balance=True):
        XX = np.ascontiguousarray(np.asarray(XX, dtype=np.float64))
        if XX.ndim != 2 or XX.shape[1] not in (2, 3):
            raise ValueError("points must have shape (N, 2) or (N, 3)")
        if XX.shape[0] == 0:
            raise ValueError("need at least one point")
        if leaf_size < 1:
            raise ValueError("leaf_size must be positive")
        self.XX = XX
        self.N, self.ndim = XX.shape
        self.leaf_size = int(leaf_size)
        d = self.ndim
        self._nchild = 2 ** d
        self._D = 30 if d == 2 else 20

        lo = XX.min(axis=0)
        hi = XX.max(axis=0)
        ext = float(np.max(hi - lo))
        if ext <= 0.0:
            ext = 1.0
        self.L0 = ext * (1.0 + 1e-12)
        self.c0 = 0.5 * (lo + hi)
        self._lo = self.c0 - 0.5 * self.L0

        scale = float(2 ** self._D)
        q = np.floor((XX - self._lo) / self.L0 * scale).astype(np.int64)
        np.clip(q, 0, 2 ** self._D - 1, out=q)
        codes = morton_encode(q, self._D)
        self._perm = np.argsort(codes, kind="stable")
        self._codes = codes[self._perm]

        self._build(balance)

    # ------------------------------------------------------------------ build
    def _range_of(self, level, codes):
        shift = d_shift = self.ndim * (self._D - level)
        del d_shift
        lo = np.searchsorted(self._codes, codes << shift, side="left")
        hi = np.searchsorted(self._codes, (codes + 1) << shift, side="left")
        return lo, hi

    def _build(self, balance):
        d = self.ndim
        nc = self._nchild
        # per level: dict code -> [start, end]; leaf status derived at the end
        levels = [{0: (0, self.N)}]
        split = [set()]  # codes of boxes at each level that have children
        lev = 0
        while True:
            to_split = [c for c, (s, e) in levels[lev].items() if e - s > self.leaf_size]
            if not to_split:
                break
            if lev + 1 > self._D:
                raise ValueError("too many coincident points for the requested leaf_size")
            parents = np.array(sorted(to_split), dtype=np.int64)
            child_codes = (parents[:, None] * nc + np.arange(nc, dtype=np.int64)[None, :]).ravel()
            s, e = self._range_of(lev + 1, child_codes)
            keep = e > s
            levels.append({int(c): (int(a), int(b)) for c, a, b in zip(child_codes[keep], s[keep], e[keep])})
            split.append(set())
            split[lev].update(int(c) for c in parents)
            lev += 1

        if balance:
            self._balance(levels, split)

        # number the boxes: level by level, Morton order inside a level
        nlevels = len(levels)
        while nlevels > 1 and len(levels[nlevels - 1]) == 0:
            nlevels -= 1
        self.nlevels = nlevels
        level_sep = np.zeros(nlevels + 1, dtype=np.int64)
        codes_lev = []
        for l in range(nlevels):
            cl = np.array(sorted(levels[l].keys()), dtype=np.int64)
            codes_lev.append(cl)
            level_sep[l + 1] = level_sep[l] + cl.shape[0]
        self.level_sep = level_sep
        self.nboxes = int(level_sep[-1])
        self._codes_lev = codes_lev

        nb = self.nboxes
        self._box_level = np.zeros(nb, dtype=np.int64)
        self._box_code = np.zeros(nb, dtype=np.int64)
        self._box_start = np.zeros(nb, dtype=np.int64)
        self._box_end = np.zeros(nb, dtype=np.int64)
        self._box_icoord = np.zeros((nb, d), dtype=np.int64)
        for l in range(nlevels):
            a, b = level_sep[l], level_sep[l + 1]
            cl = codes_lev[l]
            self._box_level[a:b] = l
            self._box_code[a:b] = cl
            se = np.array([levels[l][int(c)] for c in cl], dtype=np.int64).reshape(-1, 2)
            self._box_start[a:b] = se[:, 0]
            self._box_end[a:b] = se[:, 1]
            self._box_icoord[a:b] = morton_decode(cl, max(l, 1), d)

        children = -np.ones((nb, nc), dtype=np.int64)
        parent = -np.ones(nb, dtype=np.int64)
        for l in range(nlevels - 1):
            a, b = level_sep[l], level_sep[l + 1]
            cc = (codes_lev[l][:, None] * nc + np.arange(nc)[None, :]).ravel()
            ids = self._lookup(l + 1, cc).reshape(-1, nc)
            children[a:b] = ids
            rows = np.repeat(np.arange(a, b), nc).reshape(-1, nc)
            m = ids >= 0
            parent[ids[m]] = rows[m]
        self._children = children
        self._parent = parent
        self._is_leaf = np.all(children < 0, axis=1)
        self._leaves = np.nonzero(self._is_leaf)[0]
        self.nleaves = int(self._leaves.shape[0])

        lens = self._box_end[self._leaves] - self._box_start[self._leaves]
        if np.any(lens <= 0):
            raise AssertionError("empty leaf created")

    def _balance(self, levels, split):
        """Split leaves until touching leaves differ by at most one level."""
        d = self.ndim
        nc = self._nchild
        offsets = np.array(np.meshgrid(*([[-1, 0, 1]] * d), indexing="ij")).reshape(d, -1).T
        offsets = offsets[np.any(offsets != 0, axis=1)]

        def sorted_codes(l):
            return np.array(sorted(levels[l].keys()), dtype=np.int64)

        def split_leaf(l, code):
            s, e = levels[l][code]
            cc = code * nc + np.arange(nc, dtype=np.int64)
            if l + 1 >= len(levels):
                levels.append({})
                split.append(set())
            cs, ce = self._range_of(l + 1, cc)
            for c, a, b in zip(cc, cs, ce):
                if b > a:
                    levels[l + 1][int(c)] = (int(a), int(b))
            split[l].add(code)

        for l in range(len(levels) - 1, 1, -1):
            while True:
                cl = sorted_codes(l)
                if cl.shape[0] == 0:
                    break
                ic = morton_decode(cl, l, d)
                nbr = (ic[:, None, :] + offsets[None, :, :]).reshape(-1, d)
                ok = np.all((nbr >= 0) & (nbr < 2 ** l), axis=1)
                nbr = nbr[ok]
                to_split = set()
                # walk up: the deepest existing box containing each neighbour cell
                found = np.zeros(nbr.shape[0], dtype=bool)
                for lp in range(l, -1, -1):
                    if found.all():
                        break
                    codes_lp = morton_encode(nbr >> (l - lp), max(lp, 1)) if lp > 0 else np.zeros(nbr.shape[0], dtype=np.int64)
                    clp = sorted_codes(lp)
                    pos = np.searchsorted(clp, codes_lp)
                    pos = np.minimum(pos, max(clp.shape[0] - 1, 0))
                    exists = (clp.shape[0] > 0) & (clp[pos] == codes_lp) & ~found
                    if lp < l - 1:
                        for code in np.unique(codes_lp[exists]):
                            code = int(code)
                            if code not in split[lp]:
                                to_split.add((lp, code))
                    found |= exists
                if not to_split:
                    break
                for lp, code in sorted(to_split):
                    if code not in split[lp]:
                        split_leaf(lp, code)

    def _lookup(self, level, codes):
        """Box ids of the given Morton codes at ``level`` (``-1`` where absent)."""
        codes = np.asarray(codes, dtype=np.int64)
        out = -np.ones(codes.shape[0], dtype=np.int64)
        if level >= self.nlevels:
            return out
        cl = self._codes_lev[level]
        if cl.shape[0] == 0:
            return out
        pos = np.searchsorted(cl, codes)
        pos_c = np.minimum(pos, cl.shape[0] - 1)
        hit = cl[pos_c] == codes
        out[hit] = self.level_sep[level] + pos_c[hit]
        return out

    # ------------------------------------------------------------- accessors
    def get_leaves(self):
        return self._leaves.copy()

    def get_boxes_level(self, level):
        return np.arange(self.level_sep[level], self.level_sep[level + 1], dtype=np.int64)

    def get_box_inds(self, box):
        return self._perm[self._box_start[box]:self._box_end[box]]

    def get_box_npoints(self, boxes):
        boxes = np.asarray(boxes)
        return self._box_end[boxes] - self._box_start[boxes]

    def is_leaf(self, box):
        return bool(self._is_leaf[box])

    def are_leaves(self, boxes):
        return self._is_leaf[np.asarray(boxes, dtype=np.int64)]

    def get_box_children(self, box):
        c = self._children[box]
        return c[c >= 0]

    def get_boxes_children_fullvec(self, boxes):
        return self._children[np.asarray(boxes, dtype=np.int64)]

    def get_box_parent(self, box):
        return int(self._parent[box])

    def get_box_level(self, box):
        return int(self._box_level[box])

    def get_box_length(self, box):
        return self.L0 * 2.0 ** (-int(self._box_level[box]))

    def get_boxes_centers(self, boxes):
        boxes = np.asarray(boxes, dtype=np.int64)
        lev = self._box_level[boxes]
        h = self.L0 * 2.0 ** (-lev.astype(np.float64))
        return self._lo[None, :] + (self._box_icoord[boxes] + 0.5) * h[:, None]

    def get_box_center(self, box):
        return self.get_boxes_centers(np.array([box]))[0]

    def get_leaves_above(self, level):
        """Leaves on levels strictly coarser than ``level``."""
        return self._leaves[self._box_level[self._leaves] < level]

    def _neighbor_cells(self, boxes, include_self):
        d = self.ndim
        offsets = np.array(np.meshgrid(*([[-1, 0, 1]] * d), indexing="ij")).reshape(d, -1).T
        if not include_self:
            offsets = offsets[np.any(offsets != 0, axis=1)]
        ic = self._box_icoord[boxes]
        nbr = ic[:, None, :] + offsets[None, :, :]
        return nbr, offsets.shape[0]

    def get_boxes_colleague_neigh_fullvec(self, boxes):
        """Same-level touching boxes (self included), one row per box, padded with -1."""
        boxes = np.asarray(boxes, dtype=np.int64)
        nrow = 3 ** self.ndim
        out = -np.ones((boxes.shape[0], nrow), dtype=np.int64)
        for lev in np.unique(self._box_level[boxes]):
            sel = np.nonzero(self._box_level[boxes] == lev)[0]
            nbr, k = self._neighbor_cells(boxes[sel], True)
            valid = np.all((nbr >= 0) & (nbr < 2 ** int(lev)), axis=2)
            flat = nbr.reshape(-1, self.ndim)
            codes = morton_encode(np.clip(flat, 0, None), max(int(lev), 1)) if lev > 0 else np.zeros(flat.shape[0], np.int64)
            ids = self._lookup(int(lev), codes).reshape(-1, k)
            ids[~valid] = -1
            # compact each row: existing colleagues first, in Morton order of the offsets
            for r, row in zip(sel, ids):
                v = row[row >= 0]
                out[r, :v.shape[0]] = np.sort(v)
        return out

    def get_boxes_coarse_neigh_fullvec(self, boxes):
        """Touching leaves on coarser levels, one row per box, padded with -1."""
        boxes = np.asarray(boxes, dtype=np.int64)
        nrow = 3 ** self.ndim
        out = -np.ones((boxes.shape[0], nrow), dtype=np.int64)
        for lev in np.unique(self._box_level[boxes]):
            lev = int(lev)
            sel = np.nonzero(self._box_level[boxes] == lev)[0]
            if lev == 0:
                continue
            nbr, k = self._neighbor_cells(boxes[sel], False)
            valid = np.all((nbr >= 0) & (nbr < 2 ** lev), axis=2).reshape(-1)
            flat = nbr.reshape(-1, self.ndim)
            flat = np.clip(flat, 0, 2 ** lev - 1)
            deepest = -np.ones(flat.shape[0], dtype=np.int64)
            for lp in range(lev, -1, -1):
                todo = (deepest < 0) & valid
                if not todo.any():
                    break
                sub = flat[todo] >> (lev - lp)
                codes = morton_encode(sub, max(lp, 1)) if lp > 0 else np.zeros(sub.shape[0], np.int64)
                ids = self._lookup(lp, codes)
                idx = np.nonzero(todo)[0]
                deepest[idx[ids >= 0]] = ids[ids >= 0]
            deepest = deepest.reshape(-1, k)
            for r, row in zip(sel, deepest):
                v = row[row >= 0]
                v = v[self._is_leaf[v] & (self._box_level[v] < lev)]
                v = np.unique(v)
                out[r, :v.shape[0]] = v
        return out
