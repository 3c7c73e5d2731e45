"""Serial per-pair apply (oracle; test infrastructure only).

Restates ``RecursiveSkeletonization.matvec`` and its helpers
(``skelfmm/recursive_skel.py:784-1084``) on a build object that exposes the
ragged build data (``cand``, ``bs_list``, ``rank_list``, ``proj``,
``child_src``, the five pair lists, ``root_level``, ``skel_leaves``). Every
pair's dense block comes from ``oracle.kernels.kernel_matrix`` (exact, zero on
coincident points, as the numpy kernels the serial reference uses). Vectors are
per-box arrays; skeleton vectors are kept in the charge dtype (the reference's
``np.zeros(q_vec.shape)`` at 1006 is real-only, which would drop the imaginary
part of complex charges).
"""

from __future__ import annotations

import numpy as np

from .kernels import GEOM_SIDE, kernel_matrix


class SerialApply:
    def __init__(self, build):
        self.b = build
        self.name = build.operator.name
        self.kappa = build.kappa
        self.X = build.tree.XX
        self.G = build.geometry
        self.side = GEOM_SIDE.get(self.name)

    def _K(self, pu, pq):
        """``K(X[pu], X[pq])`` with the geometry on the operator's side (``recursive_skel.py:619-620``)."""
        gt = gs = None
        if self.side == "target":
            gt = self.G[pu]
        elif self.side == "source":
            gs = self.G[pq]
        return kernel_matrix(self.name, self.X[pu], self.X[pq], self.kappa, geom_trg=gt, geom_src=gs)

    def matvec(self, q):
        """``y = K~ q`` (``recursive_skel.py:1041-1084``); ``q`` is 1D or 2D (N, R)."""
        b = self.b
        q = np.asarray(q)
        dt = np.result_type(q.dtype, np.complex128 if self.name in ("helmholtz_2d", "helmholtz_3d",
                                                                     "helmholtz_double_layer_3d",
                                                                     "helmholtz_adjoint_double_layer_3d")
                            else np.float64)
        q = q.astype(dt)
        tail = q.shape[1:]
        nb = b.nboxes
        bs, rk = b.bs_list, b.rank_list
        # get_qorig (908-922): leaf candidates take the user's charges
        qv = [np.zeros((int(bs[i]),) + tail, dtype=dt) for i in range(nb)]
        for leaf in b.tree.get_leaves():
            qv[leaf][:] = q[b.cand[leaf]]
        qs = [np.zeros((int(bs[i]),) + tail, dtype=dt) for i in range(nb)]

        def proj_up(box):
            r = int(rk[box])
            T = b.proj[box]
            qs[box][:r] += qv[box][:r] + T @ qv[box][r:int(bs[box])]

        def proj_down(box, u, us):
            r = int(rk[box])
            T = b.proj[box]
            u[box][:r] += us[box][:r]
            u[box][r:int(bs[box])] += T.T @ us[box][:r]

        def tree_boxes(lev):
            bl = b.tree.get_boxes_level(lev)
            return bl[~b.tree.are_leaves(bl)]

        def children_to_parent(lev, vec_res, vec_mv, upward):
            # addto_skel_lev (939-966) for the tree boxes of `lev`
            for P in tree_boxes(lev):
                cs = b.child_src[P]
                for k, (c, j) in enumerate(cs):
                    if upward:
                        vec_res[P][k] += vec_mv[c][j]
                    else:
                        vec_res[c][j] += vec_mv[P][k]

        # compute_qskel (994-1017)
        for leaf in b.skel_leaves:
            proj_up(leaf)
        for lev in range(b.nlevels - 1, b.root_level, -1):
            for box in tree_boxes(lev):
                proj_up(box)
            children_to_parent(lev - 1, qv, qs, True)

        # compute_uskel (873-906)
        u = [np.zeros((int(bs[i]),) + tail, dtype=dt) for i in range(nb)]
        us = [np.zeros((int(bs[i]),) + tail, dtype=dt) for i in range(nb)]
        for U, Q in b.uskel_from_qskel_list:
            K = self._K(b.cand[U], b.cand[Q])
            rU, rQ = int(rk[U]), int(rk[Q])
            u[U] += K @ qv[Q]
            us[U][:rU] -= K[:rU, :rQ] @ qs[Q][:rQ]
        for U, Q in b.u_from_q_list:
            u[U] += self._K(b.cand[U], b.cand[Q]) @ qv[Q]
        for U, C in b.uskel_from_q_list:
            K = self._K(b.cand[U], b.cand[C])
            rU = int(rk[U])
            u[U] += K @ qv[C]
            us[U][:rU] -= K[:rU] @ qv[C]
        for C, U in b.u_from_qskel_list:
            K = self._K(b.cand[C], b.cand[U])
            rU = int(rk[U])
            u[C] += K @ qv[U]
            u[C] -= K[:, :rU] @ qs[U][:rU]

        # resultvec_from_uskel (1019-1039)
        for lev in range(b.root_level + 1, b.nlevels):
            children_to_parent(lev - 1, us, u, False)
            for box in tree_boxes(lev):
                proj_down(box, u, us)
        for leaf in b.skel_leaves:
            proj_down(leaf, u, us)

        # get_result (924-937)
        y = np.zeros(q.shape, dtype=dt)
        for leaf in b.tree.get_leaves():
            y[b.cand[leaf]] = u[leaf]
        return y
