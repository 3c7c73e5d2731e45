"""Interpolative decomposition in the reference's stacked form (oracle; test infrastructure only).

``proxy_matrix`` restates ``build_stacked_proxy_matrix`` (``skelfmm/util.py:87-124``):
symmetric point kernels stack ``[K_BP^T; K_BP^T]``; geometry-aware kernels stack
``[K_PB; K_BP^T]`` with outward proxy normals on the operator's side. Kernel
values come from the oracle's closed forms. Proxy point layouts are the build's
own definition (the reference takes them from the absent simpletree), so they
are taken from ``skelfmm_amd.proxy``.

``skeleton_id`` restates ``_skeleton_id_pivoted_qr`` (``skelfmm/util.py:13-51``).
"""

from __future__ import annotations

import numpy as np
from scipy.linalg import qr, solve_triangular

from .kernels import kernel_matrix


def proxy_points(operator, center, length, nproxy):
    from skelfmm_amd.proxy import get_cube_surface, get_sphere_surface

    side = 2.95 * length  # util.py:73
    if operator.proxy_surface == "cube":
        return get_cube_surface(center, side, nproxy)
    return get_sphere_surface(center, side, nproxy)


def proxy_matrix(operator, x_box, center, length, nproxy, geom=None, kappa=0.0):
    xp = proxy_points(operator, center, length, nproxy)
    name = operator.name
    if operator.is_symmetric and geom is None:
        Kt = kernel_matrix(name, x_box, xp, kappa).T
        return np.vstack((Kt, Kt))
    nrm = xp - np.asarray(center)[None, :]
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    side = operator.geometry_side
    K_BP = kernel_matrix(name, x_box, xp, kappa, geom_trg=geom if side == "target" else None,
                         geom_src=nrm if side == "source" else None)
    K_PB = kernel_matrix(name, xp, x_box, kappa, geom_trg=nrm if side == "target" else None,
                         geom_src=geom if side == "source" else None)
    return np.vstack((K_PB, K_BP.T))


def skeleton_id(K, rank_or_eps):
    ncols = K.shape[1]
    R, piv = qr(K, mode="r", pivoting=True)
    R = np.asarray(R)[:ncols, :]
    if rank_or_eps < 1:
        dg = np.abs(np.diag(R))
        rank = 0 if (dg.size == 0 or dg[0] == 0) else int(np.count_nonzero(dg > rank_or_eps * dg[0]))
    else:
        rank = int(rank_or_eps)
    rank = max(0, min(rank, ncols))
    idx = np.asarray(piv, dtype=int)
    if rank == 0 or rank == ncols:
        proj = np.zeros((rank, ncols - rank), dtype=K.dtype)
    else:
        proj = solve_triangular(R[:rank, :rank], R[:rank, rank:], lower=False)
    return rank, idx, proj
