"""Host build: per-level interpolative decompositions and interaction lists.

Restates the build half of ``RecursiveSkeletonization``
(``skelfmm/recursive_skel.py:19-122,142-166,209-262,326-502,504-780``) with
ragged per-box storage instead of the reference's ``(nboxes, max_bs)`` padding:

* ``cand[b]``   candidate point ids (caller's order) of box ``b``, skeleton first;
* ``rank_list[b]``, ``bs_list[b]``;
* ``proj[b]``   the interpolation matrix ``T_b`` (``rank x (bs-rank)``), with
  ``K(., cand[b][rank:]) ~= K(., cand[b][:rank]) @ T_b`` on the proxy rows;
* ``child_src[b]`` for boxes whose candidates are their children's skeletons:
  ``(child, j)`` per candidate slot, i.e. ``cand[b][k] == cand[child][j]``.

The ID itself is the reference's pivoted-QR rule (``util.py:13-51``):
``scipy.linalg.qr(K, mode='r', pivoting=True)``, rank = number of
``|R_ii| > tol*|R_00|`` (or a fixed integer rank), ``T = R11^{-1} R12``.
For symmetric point kernels the reference factors ``[K_BP^T; K_BP^T]``
(``util.py:90-101``); stacking a block on itself multiplies ``R`` by ``sqrt(2)``
and leaves the column pivots, the rank test and ``R11^{-1}R12`` unchanged, so
this build factors ``K_BP^T`` once (``stack_symmetric=True`` restores the stack).
"""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor
from time import perf_counter

import numpy as np
from scipy.linalg import qr, solve_triangular
from threadpoolctl import threadpool_limits

from . import operators as operators_mod
from .proxy import get_cube_surface, get_sphere_surface, sphere_proxy_normals

__all__ = ["RecursiveSkeletonization", "skeleton_id_pivoted_qr", "build_proxy_matrix", "nproxy_for_level"]


def skeleton_id_pivoted_qr(K, rank_or_eps):
    """Column ID of ``K`` by pivoted QR (``util.py:13-51``): ``(rank, idx, proj)``."""
    ncols = K.shape[1]
    if ncols == 0:
        return 0, np.zeros(0, dtype=np.int64), np.zeros((0, 0), dtype=K.dtype)
    R, piv = qr(K, mode="r", pivoting=True, overwrite_a=True, check_finite=False)
    R = np.asarray(R)[:ncols, :]
    if rank_or_eps < 1:
        diag = np.abs(np.diag(R))
        if diag.size == 0 or diag[0] == 0:
            rank = 0
        else:
            rank = int(np.count_nonzero(diag > rank_or_eps * diag[0]))
    else:
        rank = int(rank_or_eps)
    rank = max(0, min(rank, ncols))
    idx = np.asarray(piv, dtype=np.int64)
    if rank == 0 or rank == ncols:
        proj = np.zeros((rank, ncols - rank), dtype=K.dtype)
    else:
        proj = solve_triangular(R[:rank, :rank], R[:rank, rank:], lower=False, check_finite=False)
    return rank, idx, np.ascontiguousarray(proj)


def nproxy_for_level(ndim, tol, kappa, box_length):
    """Proxy count per level (``recursive_skel.py:73-85``)."""
    ndigits = int(-np.log10(tol / 100))
    if ndim == 2:
        nproxy = int(ndigits * 10)
        nproxy += int(4 * kappa * box_length)
    else:
        nproxy = max(int(ndigits ** 2 * 8), 400)
        nproxy += 2 * kappa * box_length * max(5, ndigits - 1) + (10 if kappa > 0 else 0)
    return int(nproxy)


def build_proxy_matrix(operator, x_box, box_center, box_length, nproxy, *, geom=None, param=0.0,
                       stack_symmetric=False):
    """Proxy matrix whose column ID skeletonizes a box (``util.py:87-124``)."""
    proxy_length = 2.95 * box_length
    if operator.proxy_surface == "cube":
        x_proxy = get_cube_surface(box_center, proxy_length, nproxy)
    elif operator.proxy_surface == "sphere":
        x_proxy = get_sphere_surface(box_center, proxy_length, nproxy)
    else:
        raise ValueError(f"Unsupported proxy surface {operator.proxy_surface!r}")

    if operator.is_symmetric and geom is None:
        K_BP = operator.numpy_matrix(x_box, x_proxy, param=param, assume_not_self_interaction=True)
        K = np.ascontiguousarray(K_BP.T)
        if stack_symmetric:
            K = np.vstack((K, K))
        return K

    normals = None
    if operator.proxy_surface == "sphere":
        normals = sphere_proxy_normals(x_proxy, box_center)
    src_geom = normals if operator.geometry_side == "source" else None
    trg_geom = normals if operator.geometry_side == "target" else None
    K_BP = operator.numpy_matrix(x_box, x_proxy, geom_trg=geom, geom_src=src_geom, param=param,
                                 assume_not_self_interaction=True)
    K_PB = operator.numpy_matrix(x_proxy, x_box, geom_trg=trg_geom, geom_src=geom, param=param,
                                 assume_not_self_interaction=True)
    return np.vstack((K_PB, K_BP.T))


def _default_workers():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


class RecursiveSkeletonization:
    """Skeletonization build over a ``BalancedTree`` (``recursive_skel.py:12-122``)."""

    def __init__(self, tree, operator, kappa, tol, max_bs=-1, max_rank=-1, geometry=None,
                 stack_symmetric=False):
        self.N = tree.N
        self.tree = tree
        self.nlevels = tree.nlevels
        self.nboxes = tree.nboxes
        self.kappa = float(kappa)
        self.operator = operators_mod._require_operator(operator)
        self.ndim = tree.ndim
        self.tol = tol
        self.geometry = None if geometry is None else np.ascontiguousarray(np.asarray(geometry, dtype=np.float64))
        self.stack_symmetric = stack_symmetric

        if self.operator.ndim > 0 and self.operator.ndim != self.ndim:
            raise ValueError("Operator dimensionality does not match the tree dimensionality")
        if self.operator.requires_geometry and self.geometry is None:
            raise ValueError("Geometry is required for this operator")
        if self.geometry is not None and self.geometry.shape[0] != self.N:
            raise ValueError("Geometry must be aligned with tree.XX")

        leaves = tree.get_leaves()
        is_leaf = tree.are_leaves(np.arange(self.nboxes))
        tree_boxes = []
        self.tree_lev_sep = np.zeros(self.nlevels + 1, dtype=np.int64)
        for lev in range(self.nlevels):
            b = tree.get_boxes_level(lev)
            b = b[~is_leaf[b]]
            tree_boxes.append(b)
            self.tree_lev_sep[lev + 1] = self.tree_lev_sep[lev] + b.shape[0]
        self.tree_boxes = np.concatenate(tree_boxes) if tree_boxes else np.zeros(0, np.int64)
        self.ntree_boxes = int(self.tree_boxes.shape[0])

        self.max_bs_leaf = tree.leaf_size
        self.nproxy_level = np.array(
            [nproxy_for_level(self.ndim, self.tol, self.kappa, tree.L0 * 2.0 ** (-lev))
             for lev in range(self.nlevels)], dtype=np.int64)
        if max_bs > 0 and max_rank > 0:
            self.max_rank = int(max_rank)
            self.max_bs = int(max(max_bs, self.max_bs_leaf))
        else:
            self.max_rank = int(self.max_bs_leaf * 0.75)
            self.max_bs = int((2 ** self.ndim) * self.max_bs_leaf)
        self.max_rank_leaf = self.max_rank

        k = self.operator.numpy_matrix(
            np.zeros((1, self.ndim)), np.zeros((1, self.ndim)),
            geom_src=None if self.geometry is None else np.zeros((1, 3)),
            geom_trg=None if self.geometry is None else np.zeros((1, 3)),
            param=self.kappa)
        self.kernel_dtype = k.dtype

        nb = self.nboxes
        self.cand = [np.zeros(0, dtype=np.int64) for _ in range(nb)]
        self.proj = [None] * nb
        self.child_src = [None] * nb
        self.bs_list = np.zeros(nb, dtype=np.int64)
        self.rank_list = np.zeros(nb, dtype=np.int64)
        self.rank_max = 0
        self.root_level = max(0, self.nlevels - 1)
        self.level_times = {}
        self.assign_leaves()

    # ------------------------------------------------------------------ leaves
    def assign_leaves(self):
        """Leaf candidates are the leaf's points (``recursive_skel.py:142-166``)."""
        self.leaf_size = 0
        for leaf in self.tree.get_leaves():
            inds = np.asarray(self.tree.get_box_inds(leaf), dtype=np.int64).copy()
            assert inds.shape[0] > 0, "Leaf box must contain points"
            if inds.shape[0] > self.max_bs_leaf:
                raise ValueError("Leaf block size exceeds maximum allowed")
            self.leaf_size = max(self.leaf_size, inds.shape[0])
            self.cand[leaf] = inds
            self.bs_list[leaf] = inds.shape[0]

    def update_lev_skel(self, lev):
        """Parent candidates = children's skeletons in child order (``recursive_skel.py:209-262``)."""
        for box in self.tree.get_boxes_level(lev):
            if self.tree.is_leaf(box):
                continue
            parts = []
            src = []
            for c in self.tree.get_box_children(box):
                r = int(self.rank_list[c])
                parts.append(self.cand[c][:r])
                src.append(np.stack([np.full(r, c, dtype=np.int64), np.arange(r, dtype=np.int64)], axis=1))
            self.cand[box] = np.concatenate(parts) if parts else np.zeros(0, np.int64)
            self.child_src[box] = np.concatenate(src) if src else np.zeros((0, 2), np.int64)
            self.bs_list[box] = self.cand[box].shape[0]

    def _skel_one(self, box, center, length, nproxy):
        pts = self.cand[box]
        XX_B = self.tree.XX[pts]
        geom_B = None if self.geometry is None else self.geometry[pts]
        K = build_proxy_matrix(self.operator, XX_B, center, length, nproxy, geom=geom_B, param=self.kappa,
                               stack_symmetric=self.stack_symmetric)
        return skeleton_id_pivoted_qr(K, self.tol)

    def _skel_level_gpu(self, gpu, lev, boxes, centers, length, nproxy, rank_model):
        cands = [self.cand[b] for b in boxes]
        res = [None] * len(boxes)
        if rank_model is not None:
            from .learned_rank import box_features

            F = box_features(self, lev, boxes, centers, length)
            k_pred = rank_model.predict_rank(F)
            res = gpu.sketched(cands, centers, length, nproxy, k_pred, **self._sketch_opts)
            kept = [isinstance(r, tuple) and len(r) == 2 and r[0] == "kept" for r in res]
            sk = [r is not None and not isinstance(r, str) and not k for r, k in zip(res, kept)]
            self.gpu_stats["kept"] += sum(kept)
            self.gpu_stats["sketched"] += sum(sk)
            self.gpu_stats["direct"] += sum(isinstance(r, str) for r in res)
            self.gpu_stats["fallback"] += sum(r is None for r in res)
            res = [r[1] if k else (None if isinstance(r, str) else r) for r, k in zip(res, kept)]
        todo = [i for i, r in enumerate(res) if r is None]
        if todo:
            sub = gpu.adaptive([cands[i] for i in todo], centers[todo], length, nproxy)
            for i, r in zip(todo, sub):
                res[i] = r
            self.gpu_stats["adaptive"] += sum(r is not None for r in sub)
        for i, r in enumerate(res):
            if r is None:  # workspace beyond one workgroup's LDS: host ID
                res[i] = self._skel_one(boxes[i], centers[i], length, nproxy)
                self.gpu_stats["cpu"] += 1
        return res

    def set_skel_box_info(self, box, rank, idx, proj):
        """Permute the box skeleton-first and store ``T`` (``recursive_skel.py:264-324``)."""
        self.rank_list[box] = rank
        self.cand[box] = self.cand[box][idx]
        if self.child_src[box] is not None:
            self.child_src[box] = self.child_src[box][idx]
        self.proj[box] = proj

    def skel_tree(self, npoints_max=500, p="auto", verbose=False, min_parallel_items=64, nworkers=None,
                  device="cpu", rank_model=None, gpu_options=None):
        """Bottom-up skeletonization until level 2 or ``sum(bs) <= npoints_max`` (``recursive_skel.py:326-502``).

        ``device="gpu"`` runs the per-box IDs of a level as batched GPU kernels
        (``skelfmm_amd.gpu_build``): the adaptive pivoted QR, or -- with ``rank_model``
        (a trained ``learned_rank.RankMLP``) -- the sketched fixed-rank QR with the
        proxy-residual check, falling back to the adaptive QR box by box. ``gpu_options``:
        ``{"profile": bool, "sketch": {keyword arguments of GPUSkeletonizer.sketched}}``.
        """
        if nworkers is None:
            nworkers = _default_workers()
        self.root_level = max(0, self.tree.nlevels - 1)
        self.build_device = device
        # boxes by path: accepted sketches, sketches rejected (fallback), predicted incompressible
        # (direct), predicted full rank and kept whole (kept), adaptive GPU IDs (fallback + direct
        # + no model), host IDs (LDS too small)
        self.gpu_stats = {"sketched": 0, "fallback": 0, "direct": 0, "kept": 0, "adaptive": 0, "cpu": 0}
        gpu = None
        if device == "gpu" and getattr(self.operator, "user_kernel", None) is not None:
            raise ValueError("user-defined operators are built on the host (device='cpu'); "
                             "their near field runs on the GPU")
        if device == "gpu":
            from .gpu_build import GPUSkeletonizer

            gpu = GPUSkeletonizer(self.operator, self.kappa, self.tol, self.tree.XX, self.geometry)
            gpu.profile = bool((gpu_options or {}).get("profile", False))
        self.gpu_skel = gpu
        self._sketch_opts = dict((gpu_options or {}).get("sketch", {}))
        pool = None
        try:
            for lev in range(self.tree.nlevels - 1, 0, -1):
                tic = perf_counter()
                self.update_lev_skel(lev)
                boxes_lev = self.tree.get_boxes_level(lev)
                active = np.concatenate((self.tree.get_leaves_above(lev), boxes_lev))
                npoints_lev = int(np.sum(self.bs_list[active]))
                if verbose:
                    print("\t Level %2d has %8d npoints" % (lev, npoints_lev))
                if lev == 2 or npoints_lev <= npoints_max:
                    self.root_level = lev
                    break
                length = self.tree.get_box_length(boxes_lev[0])
                centers = self.tree.get_boxes_centers(boxes_lev)
                nproxy = int(self.nproxy_level[lev])
                parallel = (p is not None) and nworkers > 1 and boxes_lev.shape[0] >= min_parallel_items
                if gpu is not None and self.tol < 1:
                    res = self._skel_level_gpu(gpu, lev, boxes_lev, centers, length, nproxy, rank_model)
                elif parallel:
                    # one BLAS thread per worker: the QRs run concurrently (LAPACK releases the GIL)
                    if pool is None:
                        pool = ThreadPoolExecutor(max_workers=nworkers)
                    with threadpool_limits(1):
                        res = list(pool.map(lambda a: self._skel_one(a[0], a[1], length, nproxy),
                                            zip(boxes_lev, centers)))
                else:
                    res = [self._skel_one(b, c, length, nproxy) for b, c in zip(boxes_lev, centers)]
                for box, (rank, idx, proj) in zip(boxes_lev, res):
                    self.set_skel_box_info(int(box), rank, idx, proj)
                self.level_times[lev] = perf_counter() - tic
                if verbose:
                    print("\t Level %2d skeletonized in %6.2f s, max rank %d" %
                          (lev, self.level_times[lev], int(self.rank_list[boxes_lev].max())))
        finally:
            if pool is not None:
                pool.shutdown()
        return self

    # ------------------------------------------------------------------- lists
    def setup_lists(self, verbose=False):
        self.prune_skel()
        self.get_parallel_lists(verbose)
        return self

    def prune_skel(self):
        """Size statistics of the build (``recursive_skel.py:515-539``)."""
        self.max_bs = int(np.max(self.bs_list))
        self.max_rank = int(np.max(self.rank_list))
        self.rank_max = self.max_rank
        pos = self.rank_list[self.rank_list > 0]
        self.min_rank = int(np.min(pos)) if pos.size > 0 else 0
        leaves = self.tree.get_leaves()
        self.max_bs_leaf = int(np.max(self.bs_list[leaves]))
        self.max_rank_leaf = int(np.max(self.rank_list[leaves]))

    def get_parallel_lists(self, verbose=False):
        """The five pair lists (``recursive_skel.py:622-712``)."""
        tree = self.tree
        leaf_mask = tree.are_leaves(np.arange(self.nboxes))
        leaf_pairs, tree_pairs, coarse_pairs = [], [], []
        if hasattr(tree, "pair_lists"):  # GPU tree: the same three lists from its device rows
            lp, tp, cp = tree.pair_lists(self.root_level)
            leaf_pairs, tree_pairs, coarse_pairs = [lp], [tp], [cp]
        for lev in range(self.root_level + 1, self.nlevels) if not hasattr(tree, "pair_lists") else ():
            boxes = tree.get_boxes_level(lev)
            coll = tree.get_boxes_colleague_neigh_fullvec(boxes)
            coarse = tree.get_boxes_coarse_neigh_fullvec(boxes)
            nb = coll.shape[1]
            u = np.repeat(boxes, nb).reshape(-1, nb)
            valid = coll >= 0
            both_leaf = valid & leaf_mask[u] & leaf_mask[np.where(valid, coll, 0)]
            other = valid & ~both_leaf
            leaf_pairs.append(np.stack([u[both_leaf], coll[both_leaf]], axis=1))
            tree_pairs.append(np.stack([u[other], coll[other]], axis=1))
            cv = coarse >= 0
            uc = np.repeat(boxes, coarse.shape[1]).reshape(-1, coarse.shape[1])
            coarse_pairs.append(np.stack([uc[cv], coarse[cv]], axis=1))

        def cat(x):
            x = [a for a in x if a.shape[0] > 0]
            return np.concatenate(x).astype(np.int64) if x else np.zeros((0, 2), np.int64)

        self.leaf_uskel_from_qskel_list = cat(leaf_pairs)
        self.tree_uskel_from_qskel_list = cat(tree_pairs)
        self.uskel_from_qskel_list = np.vstack((self.leaf_uskel_from_qskel_list, self.tree_uskel_from_qskel_list))
        self.uskel_from_q_list = cat(coarse_pairs)
        self.u_from_qskel_list = self.uskel_from_q_list[:, ::-1].copy()

        boxes_level = tree.get_boxes_level(self.root_level)
        leaves_above = tree.get_leaves_above(self.root_level)
        self.root_boxes = np.concatenate((leaves_above, boxes_level)).astype(np.int64)
        leaves = tree.get_leaves()
        self.skel_leaves = np.setdiff1d(leaves, tree.get_leaves_above(self.root_level + 1))
        rb = self.root_boxes
        self.u_from_q_list = np.stack([np.repeat(rb, rb.shape[0]), np.tile(rb, rb.shape[0])], axis=1)
        if verbose:
            print("\t %d root boxes on level %d, %d active points" %
                  (rb.shape[0], self.root_level, int(np.sum(self.bs_list[rb]))))

    # ---------------------------------------------------------------- helpers
    def skel_boxes_level(self, lev):
        """Skeletonized boxes (tree boxes and leaves) on level ``lev`` > root."""
        if lev <= self.root_level:
            return np.zeros(0, np.int64)
        return self.tree.get_boxes_level(lev)

    def gathered_boxes_level(self, lev):
        """Boxes on ``lev`` whose candidates are their children's skeletons."""
        if lev < self.root_level:
            return np.zeros(0, np.int64)
        b = self.tree.get_boxes_level(lev)
        return b[~self.tree.are_leaves(b)]

    @property
    def nbytes_proj(self):
        return int(sum(p.nbytes for p in self.proj if p is not None))
