"""The adaptive 2:1-balanced tree and its neighbour and pair lists, built on the GPU.

``GPUBalancedTree(XX, leaf_size)`` is the tree of ``skelfmm_amd.tree.BalancedTree`` (the
contract of the reference's ``BalancedTree``, ``skelfmm/skel_fmm.py:150``) built by the HIP
kernels of ``csrc/skelfmm_tree.hip``; every integer table equals the host statement's bit for
bit (``tests/test_gpu_tree.py``). Steps, each on the device:

1. point keys: the points' bounding box (a min/max reduction), then the Morton key of every
   point at depth D (30 bits per axis in 2D, 20 in 3D) with the host statement's FP64
   arithmetic;
2. a stable radix sort of (key, point index): every box owns a contiguous range of it;
3. levels top down: the boxes holding more than ``leaf_size`` points split, each child's range
   found by binary search in its parent's range, empty children dropped;
4. 2:1 balance, deepest level first: the kernel marks the leaves that a touching box more than
   one level deeper forces to split; they split and the level is checked again;
5. box tables (level, integer coordinates, children, parent), colleague rows (same level,
   touching, self included) and coarse rows (touching leaves on coarser levels) for every box;
6. on request (``pair_lists``), the leaf, tree and coarse pair lists of the levels below the root
   (``skelfmm/recursive_skel.py:622-712``), counted and filled per box row.

The host receives the tables once (sorted point order, box tables, neighbour rows), so every
``BalancedTree`` query answers from them. Per level the host reads one count back to size the
next level's arrays.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import hip
from .tree import BalancedTree

__all__ = ["GPUBalancedTree"]


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class GPUBalancedTree(BalancedTree):
    """``BalancedTree`` whose construction, neighbour rows and pair lists run on the GPU."""

    def __init__(self, XX, leaf_size=200, balance=True, device=None):
        XX = np.ascontiguousarray(np.asarray(XX, dtype=np.float64))
        if XX.ndim != 2 or XX.shape[1] not in (2, 3):
            raise ValueError("points must have shape (N, 2) or (N, 3)")
        if XX.shape[0] == 0:
            raise ValueError("need at least one point")
        if leaf_size < 1:
            raise ValueError("leaf_size must be positive")
        if XX.shape[0] >= 2 ** 31:
            raise ValueError("the GPU tree indexes points with 32-bit integers")
        hip.require_gpu()
        self._lib = hip.lib()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.XX = XX
        self.N, self.ndim = XX.shape
        self.leaf_size = int(leaf_size)
        d = self.ndim
        self._nchild = 2 ** d
        self._D = 30 if d == 2 else 20
        self._stream = hip.stream_handle(self.device)

        pts = torch.as_tensor(XX, device=self.device)
        mn, mx = (t.cpu().numpy() for t in torch.aminmax(pts, dim=0))
        ext = float(np.max(mx - mn))
        if ext <= 0.0:
            ext = 1.0
        self.L0 = ext * (1.0 + 1e-12)
        self.c0 = 0.5 * (mn + mx)
        self._lo = self.c0 - 0.5 * self.L0

        keys, perm = self._sorted_keys(pts)
        del pts
        codes, starts, ends, split = self._levels(keys)
        if balance:
            self._balance(keys, codes, starts, ends, split)
        self._finish(keys, perm, codes, starts, ends)

    # ------------------------------------------------------------ device steps
    def _sorted_keys(self, pts):
        L, n, dev = self._lib, self.N, self.device
        keys = torch.empty(n, dtype=torch.int64, device=dev)
        idx = torch.empty(n, dtype=torch.int32, device=dev)
        lo = (ctypes.c_double * 3)(*(list(self._lo) + [0.0] * (3 - self.ndim)))
        hip.check_tree(L.skf_tree_keys(_p(pts), n, self.ndim, lo, float(self.L0), self._D, _p(keys), _p(idx),
                                       self._stream))
        skeys = torch.empty_like(keys)
        perm = torch.empty_like(idx)
        nbytes = ctypes.c_ulonglong(0)
        end_bit = self.ndim * self._D
        hip.check_tree(L.skf_tree_sort_pairs(None, ctypes.byref(nbytes), _p(keys), _p(skeys), _p(idx), _p(perm), n,
                                             end_bit, self._stream))
        temp = torch.empty(max(int(nbytes.value), 1), dtype=torch.uint8, device=dev)
        hip.check_tree(L.skf_tree_sort_pairs(_p(temp), ctypes.byref(nbytes), _p(keys), _p(skeys), _p(idx), _p(perm), n,
                                             end_bit, self._stream))
        return skeys, perm

    def _children_of(self, keys, child_level, p_code, p_start, p_end):
        """Non-empty children (code, start, end) of the given boxes, in code order."""
        nc = self._nchild
        npar = int(p_code.shape[0])
        dev = self.device
        c_code = torch.empty(npar * nc, dtype=torch.int64, device=dev)
        c_start = torch.empty(npar * nc, dtype=torch.int32, device=dev)
        c_end = torch.empty(npar * nc, dtype=torch.int32, device=dev)
        hip.check_tree(self._lib.skf_tree_child_ranges(_p(keys), self.ndim, self._D, child_level, npar, _p(p_code),
                                                       _p(p_start), _p(p_end), _p(c_code), _p(c_start), _p(c_end),
                                                       self._stream))
        keep = c_end > c_start
        return c_code[keep], c_start[keep], c_end[keep]

    def _levels(self, keys):
        """Levels top down: boxes with more than leaf_size points split (their non-empty children)."""
        dev = self.device
        codes = [torch.zeros(1, dtype=torch.int64, device=dev)]
        starts = [torch.zeros(1, dtype=torch.int32, device=dev)]
        ends = [torch.full((1,), self.N, dtype=torch.int32, device=dev)]
        split = [torch.zeros(1, dtype=torch.int32, device=dev)]
        lev = 0
        while True:
            sel = torch.nonzero(ends[lev] - starts[lev] > self.leaf_size).flatten()
            if sel.numel() == 0:
                break
            if lev + 1 > self._D:
                raise ValueError("too many coincident points for the requested leaf_size")
            c, s, e = self._children_of(keys, lev + 1, codes[lev][sel], starts[lev][sel], ends[lev][sel])
            split[lev][sel] = 1
            codes.append(c)
            starts.append(s)
            ends.append(e)
            split.append(torch.zeros(c.shape[0], dtype=torch.int32, device=dev))
            lev += 1
        return codes, starts, ends, split

    @staticmethod
    def _concat(codes):
        sizes = np.array([c.shape[0] for c in codes], dtype=np.int64)
        sep = np.concatenate(([0], np.cumsum(sizes)))
        return torch.cat(codes), sep

    def _balance(self, keys, codes, starts, ends, split):
        """Split leaves until touching leaves differ by at most one level (deepest level first)."""
        dev = self.device
        for l in range(len(codes) - 1, 1, -1):
            while True:
                cat, sep = self._concat(codes)
                sep_dev = torch.as_tensor(sep, device=dev)
                mark = torch.zeros(cat.shape[0], dtype=torch.int32, device=dev)
                split_cat = torch.cat(split)
                hip.check_tree(self._lib.skf_tree_balance_marks(self.ndim, l, int(sep[l + 1] - sep[l]), _p(cat),
                                                                _p(sep_dev), _p(split_cat), _p(mark), self._stream))
                marked = torch.nonzero(mark).flatten().cpu().numpy()
                if marked.size == 0:
                    break
                lps = np.searchsorted(sep, marked, side="right") - 1
                # deepest level first: splitting level lp inserts boxes into level lp + 1, so the
                # marked indices of every level not yet processed stay valid
                for lp in np.unique(lps)[::-1]:
                    lp = int(lp)
                    loc = torch.as_tensor(marked[lps == lp] - sep[lp], device=dev)
                    c, s, e = self._children_of(keys, lp + 1, codes[lp][loc], starts[lp][loc], ends[lp][loc])
                    split[lp][loc] = 1
                    merged = torch.cat((codes[lp + 1], c))
                    merged, order = torch.sort(merged)
                    codes[lp + 1] = merged
                    starts[lp + 1] = torch.cat((starts[lp + 1], s))[order]
                    ends[lp + 1] = torch.cat((ends[lp + 1], e))[order]
                    split[lp + 1] = torch.cat((split[lp + 1], torch.zeros_like(s)))[order]

    def _finish(self, keys, perm, codes, starts, ends):
        """Box tables and neighbour rows on the device, then their host copies."""
        L, d, nc, dev = self._lib, self.ndim, self._nchild, self.device
        cat, sep = self._concat(codes)
        nb = int(cat.shape[0])
        self.nlevels = len(codes)
        self.level_sep = sep
        self.nboxes = nb
        sep_dev = torch.as_tensor(sep, device=dev)
        level = torch.empty(nb, dtype=torch.int32, device=dev)
        icoord = torch.empty(nb * d, dtype=torch.int32, device=dev)
        children = torch.empty(nb * nc, dtype=torch.int64, device=dev)
        parent = torch.full((nb,), -1, dtype=torch.int64, device=dev)
        hip.check_tree(L.skf_tree_box_tables(d, self.nlevels, nb, _p(cat), _p(sep_dev), _p(level), _p(icoord),
                                             _p(children), _p(parent), self._stream))
        is_leaf = (children.view(nb, nc) < 0).all(dim=1).to(torch.uint8)
        width = 3 ** d
        coll = torch.empty(nb * width, dtype=torch.int64, device=dev)
        coarse = torch.empty(nb * width, dtype=torch.int64, device=dev)
        hip.check_tree(L.skf_tree_colleagues(d, _p(cat), _p(sep_dev), _p(level), 0, nb, _p(coll), self._stream))
        hip.check_tree(L.skf_tree_coarse(d, _p(cat), _p(sep_dev), _p(level), _p(is_leaf), 0, nb, _p(coarse),
                                         self._stream))
        self._dev = {"codes": cat, "sep": sep_dev, "level": level, "is_leaf": is_leaf, "coll": coll,
                     "coarse": coarse}

        self._perm = perm.to(torch.int64).cpu().numpy()
        self._codes = keys.cpu().numpy()
        codes_h = cat.cpu().numpy()
        self._codes_lev = [codes_h[sep[l]:sep[l + 1]] for l in range(self.nlevels)]
        self._box_level = level.to(torch.int64).cpu().numpy()
        self._box_code = codes_h
        self._box_start = torch.cat(starts).to(torch.int64).cpu().numpy()
        self._box_end = torch.cat(ends).to(torch.int64).cpu().numpy()
        self._box_icoord = icoord.view(nb, d).to(torch.int64).cpu().numpy()
        self._children = children.view(nb, nc).cpu().numpy()
        self._parent = parent.cpu().numpy()
        self._is_leaf = is_leaf.cpu().numpy().astype(bool)
        self._leaves = np.nonzero(self._is_leaf)[0]
        self.nleaves = int(self._leaves.shape[0])
        self._coll = coll.view(nb, width).cpu().numpy()
        self._coarse = coarse.view(nb, width).cpu().numpy()
        lens = self._box_end[self._leaves] - self._box_start[self._leaves]
        if np.any(lens <= 0):
            raise AssertionError("empty leaf created")

    # ------------------------------------------------------------- neighbours
    def get_boxes_colleague_neigh_fullvec(self, boxes):
        """Same-level touching boxes (self included), one row per box, padded with -1."""
        return self._coll[np.asarray(boxes, dtype=np.int64)].copy()

    def get_boxes_coarse_neigh_fullvec(self, boxes):
        """Touching leaves on coarser levels, one row per box, padded with -1."""
        return self._coarse[np.asarray(boxes, dtype=np.int64)].copy()

    def pair_lists(self, root_level):
        """(leaf pairs, tree pairs, coarse pairs) of the boxes below ``root_level``, each (n, 2)
        int64: colleague pairs with both boxes leaves, the other colleague pairs, and (box,
        coarse neighbour) pairs, in box order and row order (``recursive_skel.py:622-712``)."""
        b0 = int(self.level_sep[root_level + 1]) if root_level + 1 < self.nlevels else self.nboxes
        b1 = self.nboxes
        empty = np.zeros((0, 2), np.int64)
        if b1 <= b0:
            return empty, empty.copy(), empty.copy()
        L, dev, D = self._lib, self.device, self._dev
        width = 3 ** self.ndim
        coll = D["coll"][b0 * width:]
        coarse = D["coarse"][b0 * width:]
        nrow = b1 - b0
        counts = torch.empty(nrow * 3, dtype=torch.int64, device=dev)
        hip.check_tree(L.skf_tree_pairs(width, _p(coll), _p(coarse), _p(D["is_leaf"]), b0, b1, 0, _p(counts), None,
                                        None, None, None, self._stream))
        c = counts.view(nrow, 3)
        offs = (torch.cumsum(c, dim=0) - c).contiguous()
        tot = c.sum(dim=0).cpu().numpy()
        outs = [torch.empty(max(int(t), 1) * 2, dtype=torch.int64, device=dev) for t in tot]
        hip.check_tree(L.skf_tree_pairs(width, _p(coll), _p(coarse), _p(D["is_leaf"]), b0, b1, 1, None, _p(offs),
                                        _p(outs[0]), _p(outs[1]), _p(outs[2]), self._stream))
        return tuple(o[:2 * int(t)].view(-1, 2).cpu().numpy() for o, t in zip(outs, tot))
