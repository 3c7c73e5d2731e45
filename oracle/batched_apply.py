"""Padded, chunked torch-CPU apply (oracle and CPU baseline; test infrastructure only).

Restates ``SkelFMMBatched`` (``skelfmm/batched_fmm.py:92-564``) on the CPU in
FP64 (or FP32/complex64 with ``use_single_precision``), deliberately faithful
rather than fast:

* the build is re-laid out as the reference's padded tensors: ``XX_list`` /
  ``XX_rank_list`` ``(nboxes, max_bs, d)`` with random padding coordinates
  (122-144), ``proj_list`` / ``proj_list_tree`` cropped with the ``min_rank``
  column offset (``recursive_skel.py:515-539``), ``short/long_vec_inds`` and
  ``box/child_skel_inds`` (``recursive_skel.py:714-780``);
* translations materialise dense kernel blocks per pair chunk, chunked by
  ``bytes_budget / (kernel_const * max_bs_loc^2 * elsize)`` pairs with a floor of
  50 (3D) / 500 (2D) pairs (314-372, ``util_batched.py:14-39``: 10 GB on CPU),
  self and non-self pairs split;
* scatter with ``index_add_`` (248, 299-302).

Vectors carry a trailing right-hand-side dimension (the reference's grouped
``G`` axis, 223-248), so ``matvec`` takes ``(N,)`` or ``(N, R)``.
"""

from __future__ import annotations

import numpy as np
import torch

from .kernels import COMPLEX_OPS, GEOM_SIDE, kernel_matrix_torch

KERNEL_CONST = {(2, False): 1.05, (2, True): 2.05, (3, False): 2.55, (3, True): 2.55}  # batched_fmm.py:50-55


class BatchedApplyCPU:
    def __init__(self, build, use_single_precision=False, bytes_budget=10e9, seed=0, pair_fraction=None):
        b = build
        self.b = b
        self.name = b.operator.name
        self.kappa = float(b.kappa)
        self.ndim = b.ndim
        self.nboxes = nb = b.nboxes
        self.N = b.N
        self.root_level = b.root_level
        self.nlevels = b.nlevels
        self.bytes_budget = bytes_budget
        self.side = GEOM_SIDE.get(self.name)
        self.real = torch.float32 if use_single_precision else torch.float64
        self.cplx = torch.complex64 if use_single_precision else torch.complex128
        self.kernel_const = KERNEL_CONST[(self.ndim, self.kappa > 0)]
        rng = np.random.RandomState(seed)

        bs = b.bs_list.astype(np.int64)
        rk = b.rank_list.astype(np.int64)
        tree = b.tree
        leaves = tree.get_leaves()
        is_leaf = tree.are_leaves(np.arange(nb))
        # prune_skel (recursive_skel.py:515-539)
        self.max_bs = max_bs = int(bs.max())
        self.max_rank = max_rank = int(rk.max())
        pos = rk[rk > 0]
        self.min_rank = min_rank = int(pos.min()) if pos.size else 0
        self.max_bs_leaf = max_bs_leaf = int(bs[leaves].max())
        self.max_rank_leaf = max_rank_leaf = int(rk[leaves].max())
        self.bs = torch.as_tensor(bs)
        self.rk = torch.as_tensor(rk)

        X = tree.XX
        XXl = rng.rand(nb, max_bs, self.ndim)
        XXr = rng.rand(nb, max_bs, self.ndim)
        G = b.geometry
        Gl = Gr = None
        if G is not None:
            Gl = np.zeros((nb, max_bs, 3))
            Gr = np.zeros((nb, max_bs, 3))
        for box in range(nb):
            if bs[box] > 0:
                XXl[box, :bs[box]] = X[b.cand[box]]
                XXr[box, :rk[box]] = X[b.cand[box][:rk[box]]]
                if G is not None:
                    Gl[box, :bs[box]] = G[b.cand[box]]
                    Gr[box, :rk[box]] = G[b.cand[box][:rk[box]]]
        self.XX_list = torch.as_tensor(XXl).to(self.real)
        self.XX_rank_list = torch.as_tensor(XXr).to(self.real)
        self.geom_list = None if Gl is None else torch.as_tensor(Gl).to(self.real)
        self.geom_rank_list = None if Gr is None else torch.as_tensor(Gr).to(self.real)

        kdt = self.cplx if self.name in COMPLEX_OPS else self.real
        self.vec_dtype = kdt
        tb = []
        lev_sep = [0]
        for lev in range(self.nlevels):
            bl = tree.get_boxes_level(lev)
            bl = bl[~is_leaf[bl]]
            tb.append(bl)
            lev_sep.append(lev_sep[-1] + bl.shape[0])
        self.tree_boxes = torch.as_tensor(np.concatenate(tb) if tb else np.zeros(0, np.int64))
        self.tree_lev_sep = lev_sep
        pdt = np.complex128 if self.name in COMPLEX_OPS else np.float64
        P = np.zeros((nb, max_rank_leaf, max(max_bs_leaf - min_rank, 0)), dtype=pdt)
        PT = np.zeros((len(self.tree_boxes), max_rank, max(max_bs - min_rank, 0)), dtype=pdt)
        tb_index = {int(x): i for i, x in enumerate(self.tree_boxes.tolist())}
        for box in range(nb):
            r = int(rk[box])
            if r == 0 or b.proj[box] is None:
                continue
            if is_leaf[box]:
                P[box, :r, r - min_rank:bs[box] - min_rank] = b.proj[box]
            else:
                PT[tb_index[box], :r, r - min_rank:bs[box] - min_rank] = b.proj[box]
        cast = kdt
        self.proj_list = torch.as_tensor(P).to(cast)
        self.proj_list_tree = torch.as_tensor(PT).to(cast)

        # precompute_inds (recursive_skel.py:714-780)
        short, long_ = [], []
        for leaf in leaves:
            short.append(np.arange(bs[leaf]) + leaf * max_bs)
            long_.append(b.cand[leaf])
        self.short_vec_inds = torch.as_tensor(np.concatenate(short))
        self.long_vec_inds = torch.as_tensor(np.concatenate(long_))
        box_skel, child_skel = [], []
        lev_off = [0] * (self.nlevels + 1)
        acc = 0
        for lev in range(self.nlevels):
            lev_off[lev] = acc
            if lev >= self.root_level:
                for Pb in tb[lev]:
                    cs = b.child_src[Pb]
                    where = {(int(c), int(j)): k for k, (c, j) in enumerate(cs)}
                    for c in tree.get_box_children(Pb):
                        for j in range(int(rk[c])):
                            box_skel.append(where[(int(c), j)] + Pb * max_bs)
                            child_skel.append(j + c * max_bs)
                            acc += 1
            lev_off[lev + 1] = acc
        self.box_skel_inds = torch.as_tensor(np.array(box_skel, dtype=np.int64))
        self.child_skel_inds = torch.as_tensor(np.array(child_skel, dtype=np.int64))
        self.inds_lev_offset = lev_off
        self.skel_leaves = torch.as_tensor(b.skel_leaves)

        # pair_fraction (CPU-baseline sampling, not part of the reference): keep each pair of every
        # list with that probability; pairs_kept / pairs_total is the fraction actually run
        self.pairs_total = 0
        self.pairs_kept = 0

        def tl(x):
            x = np.asarray(x, dtype=np.int64).reshape(-1, 2)
            self.pairs_total += x.shape[0]
            if pair_fraction is not None and x.shape[0] > 0:
                keep = np.random.RandomState(seed + 1).rand(x.shape[0]) < pair_fraction
                x = x[keep]
            self.pairs_kept += x.shape[0]
            return torch.as_tensor(x)

        self.leaf_uskel_from_qskel_list = tl(b.leaf_uskel_from_qskel_list)
        self.tree_uskel_from_qskel_list = tl(b.tree_uskel_from_qskel_list)
        self.uskel_from_q_list = tl(b.uskel_from_q_list)
        self.u_from_qskel_list = tl(b.u_from_qskel_list)
        self.u_from_q_list = tl(b.u_from_q_list)
        self.interactions = 0

    # ----------------------------------------------------------- translations
    def _kernel_apply(self, vec, uq, uskel, qskel, self_bool):
        """``apply_kernel_batched_matrices`` (``batched_fmm.py:181-217``)."""
        u, q = uq[:, 0], uq[:, 1]
        mu = int(self.rk[u].max()) if uskel else int(self.bs[u].max())
        mq = int(self.rk[q].max()) if qskel else int(self.bs[q].max())
        XU = (self.XX_rank_list if uskel else self.XX_list)[u, :mu]
        XQ = (self.XX_rank_list if qskel else self.XX_list)[q, :mq]
        gu = gq = None
        if self.geom_list is not None:
            gu = (self.geom_rank_list if uskel else self.geom_list)[u, :mu]
            gq = (self.geom_rank_list if qskel else self.geom_list)[q, :mq]
        K = kernel_matrix_torch(self.name, XU, XQ, self.kappa, self_bool, geom_trg=gu, geom_src=gq)
        self.interactions += int(K.shape[0]) * int(K.shape[1]) * int(K.shape[2])
        if K.dtype != vec.dtype:
            K = K.to(vec.dtype)
        return K @ vec

    def _addto(self, uq, u_vec, q_vec, self_bool):
        """``addto_translation`` (``batched_fmm.py:219-248``)."""
        u, q = uq[:, 0], uq[:, 1]
        mu = int(self.bs[u].max())
        mq = int(self.bs[q].max())
        tmp = q_vec[q, :mq]
        res = torch.zeros((uq.shape[0], self.max_bs) + tuple(u_vec.shape[2:]), dtype=u_vec.dtype)
        res[:, :mu] += self._kernel_apply(tmp, uq, False, False, self_bool)
        u_vec.index_add_(0, u, res)

    def _subtract(self, uq, u_vec, u_skel, q_vec, q_skel, self_bool):
        """``subtractskel_translation`` with the ``zero_col_bnd`` masks (``batched_fmm.py:250-302``)."""
        u, q = uq[:, 0], uq[:, 1]
        uskel = u_skel is not None
        qskel = q_skel is not None
        mu = int(self.rk[u].max()) if uskel else int(self.bs[u].max())
        mq = int(self.rk[q].max()) if qskel else int(self.bs[q].max())
        if qskel:
            tmp = q_skel[q, :mq].clone()
            mask = torch.arange(mq)[None, :] < self.rk[q][:, None]
            tmp = tmp * mask.reshape(mask.shape + (1,) * (tmp.dim() - 2))
        else:
            tmp = q_vec[q, :mq]
        res = torch.zeros((uq.shape[0], self.max_bs) + tuple(u_vec.shape[2:]), dtype=u_vec.dtype)
        res[:, :mu] -= self._kernel_apply(tmp, uq, uskel, qskel, self_bool)
        if uskel:
            mask = torch.arange(self.max_bs)[None, :] < self.rk[u][:, None]
            res = res * mask.reshape(mask.shape + (1,) * (res.dim() - 2))
            u_skel.index_add_(0, u, res)
        else:
            u_vec.index_add_(0, u, res)

    def _translation(self, uq, u_vec, u_skel, q_vec, q_skel, max_bs_loc):
        """``translation`` (``batched_fmm.py:314-372``): self/non-self split, memory-budget chunks."""
        n = uq.shape[0]
        if n == 0:
            return
        selfb = uq[:, 0] == uq[:, 1]
        es = torch.finfo(self.real).bits // 8 * (2 if u_vec.is_complex() else 1)
        per_pair = self.kernel_const * max_bs_loc * max_bs_loc * es
        step = max(50 if self.ndim == 3 else 500, int(self.bytes_budget / max(per_pair, 1.0)))
        for sel, flag in ((~selfb, False), (selfb, True)):
            part = uq[sel]
            for a in range(0, part.shape[0], step):
                ch = part[a:a + step]
                if u_skel is None and q_skel is None:
                    self._addto(ch, u_vec, q_vec, flag)
                else:
                    self._subtract(ch, u_vec, u_skel, q_vec, q_skel, flag)

    # ----------------------------------------------------------- level passes
    def _tree_boxes(self, lev):
        return self.tree_boxes[self.tree_lev_sep[lev]:self.tree_lev_sep[lev + 1]]

    def _addto_skel_lev(self, lev, vec_res, vec_mv, upward):
        """``addto_skel_lev`` (``batched_fmm.py:381-400``)."""
        if lev < 0:
            return
        a, b = self.inds_lev_offset[lev], self.inds_lev_offset[lev + 1]
        bsi = self.box_skel_inds[a:b]
        csi = self.child_skel_inds[a:b]
        tail = tuple(vec_res.shape[2:])
        vr = vec_res.reshape((-1,) + tail)
        vm = vec_mv.reshape((-1,) + tail)
        if upward:
            vr.index_add_(0, bsi, vm[csi].contiguous())
        else:
            vr.index_add_(0, csi, vm[bsi].contiguous())

    def _proj_levtree(self, lev, vec, upward):
        """``apply_proj_levtree`` (``batched_fmm.py:402-425``)."""
        tb = self._tree_boxes(lev)
        if tb.shape[0] == 0:
            return
        P = self.proj_list_tree[self.tree_lev_sep[lev]:self.tree_lev_sep[lev + 1], :self.max_rank,
                                :self.max_bs - self.min_rank]
        v = vec if vec.dim() == 3 else vec.unsqueeze(-1)
        if upward:
            v[tb, :self.max_rank] += P @ v[tb, self.min_rank:self.max_bs]
        else:
            v[tb, self.min_rank:self.max_bs] += torch.transpose(P, -1, -2) @ v[tb, :self.max_rank]

    def _proj_leaves(self, vec, upward):
        """``apply_proj_leaves`` (``batched_fmm.py:427-438``)."""
        v = vec if vec.dim() == 3 else vec.unsqueeze(-1)
        if upward:
            v[:, :self.max_rank_leaf] += self.proj_list @ v[:, self.min_rank:self.max_bs_leaf]
        else:
            v[:, self.min_rank:self.max_bs_leaf] += torch.transpose(self.proj_list, -1, -2) @ v[:, :self.max_rank_leaf]

    def matvec(self, q):
        """``SkelFMMBatched.matvec`` (``batched_fmm.py:539-564``) for a CPU tensor or array."""
        q = torch.as_tensor(q)
        dt = self.vec_dtype
        if q.is_complex() and not dt.is_complex:
            dt = self.cplx
        q = q.to(dt)
        squeeze = q.dim() == 1
        if squeeze:
            q = q[:, None]
        self.interactions = 0
        tail = tuple(q.shape[1:])
        # get_qorig (441-447)
        qo = torch.zeros((self.nboxes * self.max_bs,) + tail, dtype=dt)
        qo.index_add_(0, self.short_vec_inds, q[self.long_vec_inds].contiguous())
        q_vec = qo.reshape((self.nboxes, self.max_bs) + tail)
        # compute_qskel (450-467)
        q_skel = torch.zeros_like(q_vec)
        q_skel[self.skel_leaves] += q_vec[self.skel_leaves]
        self._proj_leaves(q_skel, True)
        for lev in range(self.nlevels - 1, self.root_level, -1):
            tb = self._tree_boxes(lev)
            q_skel[tb] += q_vec[tb]
            self._proj_levtree(lev, q_skel, True)
            self._addto_skel_lev(lev - 1, q_vec, q_skel, True)
        u_skel = torch.zeros_like(q_vec)
        u_vec = torch.zeros_like(q_vec)
        # addto_uvec (496-511)
        lists = torch.cat((self.u_from_qskel_list, self.u_from_q_list, self.tree_uskel_from_qskel_list,
                           self.uskel_from_q_list))
        self._translation(lists, u_vec, None, q_vec, None, self.max_bs)
        self._translation(self.leaf_uskel_from_qskel_list, u_vec, None, q_vec, None, self.max_bs_leaf)
        # subtract_from_uskel (514-534)
        coll = torch.cat((self.tree_uskel_from_qskel_list, self.leaf_uskel_from_qskel_list))
        self._translation(coll, u_vec, u_skel, q_vec, q_skel, self.max_rank)
        self._translation(self.uskel_from_q_list, u_vec, u_skel, q_vec, None, self.max_bs)
        self._translation(self.u_from_qskel_list, u_vec, None, q_vec, q_skel, self.max_bs)
        # resultvec_from_uskel (469-483)
        for lev in range(self.root_level, self.nlevels):
            tb = self._tree_boxes(lev)
            self._addto_skel_lev(lev - 1, u_skel, u_vec, False)
            self._proj_levtree(lev, u_skel, False)
            u_vec[tb] += u_skel[tb]
        self._proj_leaves(u_skel, False)
        u_vec[self.skel_leaves] += u_skel[self.skel_leaves]
        # get_result (486-490)
        res = torch.zeros((self.N,) + tail, dtype=dt)
        flat = u_vec.reshape((-1,) + tail)
        res.index_add_(0, self.long_vec_inds, flat[self.short_vec_inds].contiguous())
        return res[:, 0] if squeeze else res
