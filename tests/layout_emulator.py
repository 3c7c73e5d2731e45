"""Numpy emulation of ``skf_apply`` over a ``BoxLayout`` (test helper).

Executes exactly the index maps and work lists the HIP kernels consume
(charge gather, upward gather + GEMV per stage, near-field work items with the
SRC_NEG / SRC_SELF flags, downward parent step + GEMV^T, result scatter), with
dense blocks from the oracle's clamped torch kernels. Agreement with the
oracle's padded apply validates the layout on the CPU; agreement of the GPU
with the oracle then isolates the kernels.
"""

import numpy as np
import torch

from oracle.kernels import GEOM_SIDE, kernel_matrix_torch


def emulate_apply(L, op_name, kappa, x):
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
    for g, t0, sb, se in L.items:
        nt = min(64, int(L.g_n[g]) - t0)
        ts = L.g_pt[g] + t0 + np.arange(nt)
        acc = np.zeros(nt, dtype=dt)
        for s in range(sb, se):
            ss = L.e_pt[s] + np.arange(L.e_n[s])
            XT = torch.as_tensor(P[:, ts].T)[None]
            XS = torch.as_tensor(P[:, ss].T)[None]
            gt = torch.as_tensor(G[:, ts].T)[None] if side == "target" else None
            gs = torch.as_tensor(G[:, ss].T)[None] if side == "source" else None
            K = kernel_matrix_torch(op_name, XT, XS, kappa, False, geom_trg=gt, geom_src=gs)[0].numpy()
            if L.e_flags[s] & 2:
                li = t0 + np.arange(nt)
                lj = np.arange(L.e_n[s])
                K = np.where(li[:, None] == lj[None, :], 0.0, K)
            v = K @ q[L.e_q[s] + np.arange(L.e_n[s])]
            acc += -v if (L.e_flags[s] & 1) else v
        u[L.g_out[g] + t0 + np.arange(nt)] += acc

    for S in L.down:
        us = u[ns + S.gather_dst] + u[S.gather_src]
        u[ns + S.gather_dst] = us
        u[S.gather_dst] += us
        for off, bs, r, toff in zip(S.box_off, S.box_bs, S.box_rank, S.box_toff):
            u[off + r:off + bs] += Tmat(toff, r, bs).T @ u[ns + off:ns + off + r]

    y = np.zeros(x.shape[0], dtype=dt)
    y[L.leaf_user] = u[L.leaf_slot]
    return y
