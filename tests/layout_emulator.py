"""Numpy emulation of ``skf_apply`` over a ``BoxLayout`` (test helper).

Executes exactly the index maps and work lists the HIP kernels consume
(charge gather, upward gather + GEMV per stage, near-field work items with the
SRC_NEG / SRC_SELF flags -- or the symmetric split: one-sided rest plus pair entries
evaluated in both directions, K(s, t) for the reverse one -- downward parent step + GEMV^T, result scatter), with
dense blocks from the oracle's clamped torch kernels. Agreement with the
oracle's padded apply validates the layout on the CPU; agreement of the GPU
with the oracle then isolates the kernels.
"""

import numpy as np
import torch

from oracle.kernels import GEOM_SIDE, kernel_matrix_torch


def _near(lists, tpi, P, G, side, op_name, kappa, q, u, dt, symmetric):
    """Every work item with all of its source entries at once (one kernel block per item)."""
    e_pt = np.asarray(lists.e_pt, np.int64)
    e_n = np.asarray(lists.e_n, np.int64)
    e_q = np.asarray(lists.e_q, np.int64)
    e_fl = np.asarray(lists.e_flags, np.int64)
    for g, t0, sb, se in lists.items:
        nt = min(tpi, int(lists.g_n[g]) - t0)
        tl = t0 + np.arange(nt)
        ts = lists.g_pt[g] + tl
        ents = np.arange(sb, se)
        cnt = e_n[ents]
        first = np.repeat(np.cumsum(cnt) - cnt, cnt)
        lj = np.arange(int(cnt.sum())) - first  # local index inside the entry
        ent = np.repeat(ents, cnt)
        ss = e_pt[ent] + lj
        jq = e_q[ent] + lj
        sgn = np.where(e_fl[ent] & 1, -1.0, 1.0)
        XT = torch.as_tensor(P[:, ts].T)[None]
        XS = torch.as_tensor(P[:, ss].T)[None]
        gt = torch.as_tensor(G[:, ts].T)[None] if side == "target" else None
        gs = torch.as_tensor(G[:, ss].T)[None] if side == "source" else None
        K = kernel_matrix_torch(op_name, XT, XS, kappa, False, geom_trg=gt, geom_src=gs)[0].numpy()
        selfcol = (e_fl[ent] & 2) != 0
        if selfcol.any():
            assert not symmetric, "self segments stay one-sided"
            K = np.where(selfcol[None, :] & (tl[:, None] == lj[None, :]), 0.0, K)
        K = K * sgn[None, :]
        u[lists.g_out[g] + tl] += K @ q[jq]
        if symmetric:  # reverse direction: potentials of the entries at their charge slots
            if side is None:
                Kr = K.T  # single layers: K(s, t) = K(t, s)
            else:  # double layers: K(s, t), the two sets' roles and normals exchanged
                gt2 = torch.as_tensor(G[:, ss].T)[None] if side == "target" else None
                gs2 = torch.as_tensor(G[:, ts].T)[None] if side == "source" else None
                Kr = kernel_matrix_torch(op_name, XS, XT, kappa, False, geom_trg=gt2, geom_src=gs2)[0].numpy()
                Kr = Kr * sgn[:, None]
            np.add.at(u, jq, Kr @ q[lists.g_out[g] + tl])


def emulate_apply(L, op_name, kappa, x, split=False):
    x = np.asarray(x)
    cplx = np.iscomplexobj(x) or op_name in ("helmholtz_2d", "helmholtz_3d", "helmholtz_double_layer_3d",
                                             "helmholtz_adjoint_double_layer_3d")
    dt = np.complex128 if cplx else np.float64
    ns = L.nslots
    q = np.zeros(2 * ns, dtype=dt)
    u = np.zeros(2 * ns, dtype=dt)
    q[L.leaf_slot] = x[L.leaf_user]

    def Tmat(toff, r, bs):
        return L.T[toff:toff + r * (bs - r)].reshape(r, bs - r)

    for S in L.up:
        q[S.gather_dst] = q[ns + S.gather_src]
        for off, bs, r, toff in zip(S.box_off, S.box_bs, S.box_rank, S.box_toff):
            q[ns + off:ns + off + r] = q[off:off + r] + Tmat(toff, r, bs) @ q[off + r:off + bs]

    P = L.points
    G = L.normals
    side = GEOM_SIDE.get(op_name)
    if split:
        assert L.split is not None
        _near(L.split[0], 64, P, G, side, op_name, kappa, q, u, dt, False)
        _near(L.split[1], L.split[1].tpi, P, G, side, op_name, kappa, q, u, dt, True)
    else:
        _near(L, 64, P, G, side, op_name, kappa, q, u, dt, False)

    for S in L.down:
        us = u[ns + S.gather_dst] + u[S.gather_src]
        u[ns + S.gather_dst] = us
        u[S.gather_dst] += us
        for off, bs, r, toff in zip(S.box_off, S.box_bs, S.box_rank, S.box_toff):
            u[off + r:off + bs] += Tmat(toff, r, bs).T @ u[ns + off:ns + off + r]

    y = np.zeros(x.shape[0], dtype=dt)
    y[L.leaf_user] = u[L.leaf_slot]
    return y
