"""Near-field work tables of the box layout formed on the GPU.

The same tables as ``layout._near_field`` (groups, entries and work items of the five
interaction lists of ``skelfmm/recursive_skel.py:622-712``, applied as in
``skelfmm/batched_fmm.py:496-534``, and their symmetric split), built with device sorts, run
boundaries and prefix sums instead of host numpy; the host receives the finished tables.
Every array equals the host statement's bit for bit (``tests/test_gpu_tree.py``):

* rows (target box, target kind, source box, source kind, charge kind, subtract) of every
  list, rows with an empty target or source set dropped;
* the symmetric split (kernels of ``layout.SYMMETRIC_KERNEL_IDS``): a row pairs with its mirror when both are unique
  and each charge kind equals its source set's kind; the representative is the row whose target
  set pads less to the symmetric kernel's 64-lane rows (``layout.pair_representative``);
* per table: rows stably sorted by (target box, target kind, source box), one group per target
  set, one entry per row, work items of up to ``tpi`` targets and a source-run window of
  ``max_sources`` points, heaviest first (stable).
"""

from __future__ import annotations

import numpy as np
import torch

from .layout import (SRC_NEG, SRC_SELF, SYM_TARGETS_PER_ITEM, SYMMETRIC_KERNEL_IDS, TARGETS_PER_ITEM, NearLists,
                     pair_representative)

__all__ = ["near_field_tables"]


def _ranges(counts):
    """Concatenation of arange(c) for c in counts (device)."""
    tot = int(counts.sum())
    if tot == 0:
        return torch.zeros(0, dtype=torch.int64, device=counts.device)
    starts = torch.repeat_interleave(torch.cumsum(counts, 0) - counts, counts)
    return torch.arange(tot, dtype=torch.int64, device=counts.device) - starts


def _stable_order(key):
    return torch.sort(key, stable=True).indices


def _items(g_n, g_e0, g_e1, e_n, max_sources, tpi):
    """Work items {group, first target, first entry, end entry} and their costs, heaviest first."""
    dev = g_n.device
    ne = g_e1 - g_e0
    if g_n.shape[0] == 0 or int(ne.sum()) == 0:
        return np.zeros((0, 4), dtype=np.int32), np.zeros(0, dtype=np.int64)
    cums = torch.cat((torch.zeros(1, dtype=torch.int64, device=dev), torch.cumsum(e_n, 0)))
    gid = torch.repeat_interleave(torch.arange(g_n.shape[0], device=dev), ne)
    ent = torch.repeat_interleave(g_e0, ne) + _ranges(ne)
    loc = cums[ent] - cums[g_e0[gid]]
    bucket = torch.div(loc, int(max_sources), rounding_mode="floor")
    new_run = torch.ones(ent.shape[0], dtype=torch.bool, device=dev)
    new_run[1:] = (gid[1:] != gid[:-1]) | (bucket[1:] != bucket[:-1])
    ra = ent[new_run]
    rg = gid[new_run]
    rb = torch.cat((ra[1:], torch.zeros(1, dtype=torch.int64, device=dev)))
    last_of_group = torch.cat((rg[1:] != rg[:-1], torch.ones(1, dtype=torch.bool, device=dev)))
    rb[last_of_group] = g_e1[rg[last_of_group]]
    rsrc = cums[rb] - cums[ra]
    nblk = torch.div(g_n[rg] + tpi - 1, tpi, rounding_mode="floor")
    item_run = torch.repeat_interleave(torch.arange(ra.shape[0], device=dev), nblk)
    t0 = _ranges(nblk) * tpi
    ig = rg[item_run]
    items = torch.stack([ig, t0, ra[item_run], rb[item_run]], dim=1)
    costs = torch.clamp(g_n[ig] - t0, max=tpi) * rsrc[item_run]
    order = _stable_order(-costs)
    return (np.ascontiguousarray(items[order].to(torch.int32).cpu().numpy()),
            costs[order].cpu().numpy().astype(np.int64))


def _tables(R, off, bs, rk, nslots, max_sources, tpi, nboxes):
    """NearLists from rows R (device int64, columns tbox tkind sbox skind qkind neg)."""
    NL = NearLists()
    NL.tpi = int(tpi)
    R = R[_stable_order((R[:, 0] * 2 + R[:, 1]) * nboxes + R[:, 2])]
    scount = torch.where(R[:, 3] == 0, bs[R[:, 2]], rk[R[:, 2]])
    self_pair = (R[:, 0] == R[:, 2]) & (R[:, 1] == R[:, 3])
    key = R[:, 0] * 2 + R[:, 1]
    gkeys, gcount = torch.unique_consecutive(key, return_counts=True)
    gstart = torch.cumsum(gcount, 0) - gcount
    g_box = torch.div(gkeys, 2, rounding_mode="floor")
    g_kind = gkeys % 2
    g_n = torch.where(g_kind == 0, bs[g_box], rk[g_box])
    g_e0, g_e1 = gstart, gstart + gcount
    NL.items, NL.item_cost = _items(g_n, g_e0, g_e1, scount, max_sources, tpi)
    tn = torch.repeat_interleave(g_n, gcount)
    NL.interactions = int((tn * scount).sum())

    def host(t, dt):
        return t.cpu().numpy().astype(dt)

    NL.e_pt = host(off[R[:, 2]], np.int32)
    NL.e_n = host(scount, np.int32)
    NL.e_qkind = host(R[:, 4], np.int32)
    NL.e_q = host(off[R[:, 2]] + R[:, 4] * nslots, np.int32)
    NL.e_flags = host(R[:, 5] * SRC_NEG + self_pair.to(torch.int64) * SRC_SELF, np.int32)
    NL.e_box = host(R[:, 2], np.int32)
    NL.g_pt = host(off[g_box], np.int32)
    NL.g_n = host(g_n, np.int32)
    NL.g_out = host(off[g_box] + g_kind * nslots, np.int32)
    NL.g_box = host(g_box, np.int32)
    NL.g_kind = host(g_kind, np.int32)
    NL.g_e0 = host(g_e0, np.int64)
    NL.g_e1 = host(g_e1, np.int64)
    return NL


def _pair_rows(R, nboxes, sizes=None):
    """(one-sided rows, pair representatives) as ``layout.pair_rows``."""
    if R.shape[0] == 0:
        return R, R
    a = R[:, 0] * 2 + R[:, 1]
    b = R[:, 2] * 2 + R[:, 3]
    nb2 = 2 * int(nboxes)
    key = (a * nb2 + b) * 2 + R[:, 5]
    pkey = (b * nb2 + a) * 2 + R[:, 5]
    skey = torch.sort(key).values
    uk, cnt = torch.unique_consecutive(skey, return_counts=True)
    single = torch.isin(key, uk[cnt == 1])
    ok = single & (R[:, 0] != R[:, 2]) & (R[:, 4] == R[:, 3])
    paired = ok & torch.isin(pkey, key[ok])
    rep = paired & pair_representative(a, b, sizes)
    return R[~paired], R[rep]


def near_field_tables(fmm, off, bs, rk, nslots, max_sources, device):
    """(full, split): the near-field NearLists of build ``fmm`` and, for the single-layer kernels,
    the (one-sided, symmetric) pair of NearLists, or None."""
    dev = torch.device(device)
    nb = int(fmm.nboxes)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int64), device=dev)
    off_d, bs_d, rk_d = t(off), t(bs), t(rk)
    spec = [(fmm.u_from_qskel_list, 0, 0, 0, 0), (fmm.u_from_q_list, 0, 0, 0, 0),
            (fmm.tree_uskel_from_qskel_list, 0, 0, 0, 0), (fmm.uskel_from_q_list, 0, 0, 0, 0),
            (fmm.leaf_uskel_from_qskel_list, 0, 0, 0, 0),
            (fmm.uskel_from_qskel_list, 1, 1, 1, 1), (fmm.uskel_from_q_list, 1, 0, 0, 1),
            (fmm.u_from_qskel_list, 0, 1, 1, 1)]
    parts = []
    for lst, tk, sk, qk, neg in spec:
        if lst.shape[0] == 0:
            continue
        p = t(lst)
        consts = torch.tensor([tk, sk, qk, neg], dtype=torch.int64, device=dev).expand(p.shape[0], 4)
        parts.append(torch.stack([p[:, 0], consts[:, 0], p[:, 1], consts[:, 1], consts[:, 2], consts[:, 3]], dim=1))
    R = torch.cat(parts) if parts else torch.zeros((0, 6), dtype=torch.int64, device=dev)
    tcount = torch.where(R[:, 1] == 0, bs_d[R[:, 0]], rk_d[R[:, 0]])
    scount = torch.where(R[:, 3] == 0, bs_d[R[:, 2]], rk_d[R[:, 2]])
    keep = (tcount > 0) & (scount > 0)
    R, tcount, scount = R[keep], tcount[keep], scount[keep]
    full = _tables(R, off_d, bs_d, rk_d, nslots, max_sources, TARGETS_PER_ITEM, nb)
    split = None
    if fmm.operator.kernel_id in SYMMETRIC_KERNEL_IDS:
        one_rows, sym_rows = _pair_rows(R, nb, sizes=(tcount, scount))
        if sym_rows.shape[0]:
            one = _tables(one_rows, off_d, bs_d, rk_d, nslots, max_sources, TARGETS_PER_ITEM, nb)
            sym = _tables(sym_rows, off_d, bs_d, rk_d, nslots, max_sources, SYM_TARGETS_PER_ITEM, nb)
            assert one.interactions + 2 * sym.interactions == full.interactions
            split = (one, sym)
    return full, split
