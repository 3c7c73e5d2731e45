"""Device layout of a skeletonized build ("box layout").

Every box with candidates owns a contiguous range of *slots*; its candidate
points sit there skeleton first, so a box's skeleton is the first ``rank``
slots of its range. Vectors are ``[2*nslots][width]``: the first half is the
per-box candidate vector (the reference's ``q_vec``/``u_vec``), the second half
the skeleton vector (``q_skel``/``u_skel``) at the same slot numbers
(``skelfmm/batched_fmm.py:450-564`` uses ``(nboxes, max_bs)`` padded tensors
for the same four vectors).

Near-field work is one list for all five interaction lists of the reference
(``recursive_skel.py:622-712``, ``batched_fmm.py:496-534``). For a target box U:

* group (U, all) accumulates into ``u_vec[U]`` the ADD terms of every list
  (sources: all candidates of Q with ``q_vec``) and the SUB term of the
  coarse-down list (sources: the skeleton of the finer box with ``q_skel``);
* group (U, skel) accumulates into ``u_skel[U]`` the SUB terms of the
  colleague lists (skeleton of Q with ``q_skel``) and of the coarse-up list
  (all candidates of the coarse leaf with ``q_vec``).

Work items cover 64 targets of one group and a run of its source entries of at
most ``max_sources`` points; they are sorted by cost, heaviest first.
"""

from __future__ import annotations

import numpy as np

__all__ = ["BoxLayout", "build_layout"]

SRC_NEG = 1
SRC_SELF = 2
TARGETS_PER_ITEM = 64
ROWS_PER_ITEM = 16
COLS_PER_ITEM = 256


class LevelStage:
    """Arrays of one upward or downward interpolation stage."""

    def __init__(self, level, boxes, box_off, box_bs, box_rank, box_toff, gather_dst, gather_src, items):
        self.level = level
        self.boxes = boxes
        self.box_off = box_off
        self.box_bs = box_bs
        self.box_rank = box_rank
        self.box_toff = box_toff
        self.gather_dst = gather_dst
        self.gather_src = gather_src
        self.items = items


class BoxLayout:
    pass


def _level_tables(fmm, boxes, off, toff):
    boxes = np.asarray(boxes, dtype=np.int64)
    return (boxes,
            off[boxes].astype(np.int32),
            fmm.bs_list[boxes].astype(np.int32),
            fmm.rank_list[boxes].astype(np.int32),
            toff[boxes].astype(np.int64))


def _ranges(counts):
    """Concatenation of arange(c) for c in counts."""
    counts = np.asarray(counts, dtype=np.int64)
    tot = int(counts.sum())
    if tot == 0:
        return np.zeros(0, np.int64)
    starts = np.repeat(np.cumsum(counts) - counts, counts)
    return np.arange(tot, dtype=np.int64) - starts


def _box_items(lengths, step):
    """Items {box index, first row/column} covering lengths[i] in steps of ``step``."""
    lengths = np.asarray(lengths, dtype=np.int64)
    cnt = (lengths + step - 1) // step
    idx = np.repeat(np.arange(lengths.shape[0], dtype=np.int64), cnt)
    first = _ranges(cnt) * step
    return np.stack([idx, first], axis=1).astype(np.int32).reshape(-1, 2)


def build_layout(fmm, max_sources=16384, near_field_device=None):
    """Flatten a ``RecursiveSkeletonization`` build into device-ready arrays.

    ``near_field_device`` (a torch device) forms the near-field work tables there
    (``gpu_layout.near_field_tables``, identical to the host tables); None forms them here."""
    tree = fmm.tree
    nb = fmm.nboxes
    bs = fmm.bs_list.astype(np.int64)
    rk = fmm.rank_list.astype(np.int64)
    L = BoxLayout()
    L.ndim = fmm.ndim
    L.N = fmm.N
    L.nboxes = nb
    L.root_level = fmm.root_level

    off = -np.ones(nb, dtype=np.int64)
    has = bs > 0
    off[has] = np.concatenate(([0], np.cumsum(bs[has])[:-1])) if has.any() else []
    nslots = int(bs.sum())
    L.nslots = nslots
    L.off = off

    hb = np.nonzero(has)[0]
    slot_box = np.repeat(hb, bs[hb]).astype(np.int64)
    slot_pt = (np.concatenate([fmm.cand[b] for b in hb]) if hb.size else np.zeros(0)).astype(np.int64)
    L.slot_point = slot_pt
    L.slot_box = slot_box
    L.points = np.ascontiguousarray(tree.XX[slot_pt].T)  # (d, nslots)
    L.normals = None
    if fmm.geometry is not None:
        L.normals = np.ascontiguousarray(fmm.geometry[slot_pt].T)
    centers = np.zeros((nb, 3))
    centers[:, :fmm.ndim] = tree.get_boxes_centers(np.arange(nb))
    L.centers = centers

    # charge gather / result scatter over all leaf slots (every point exactly once)
    leaves = tree.get_leaves()
    L.leaf_slot = (np.repeat(off[leaves], bs[leaves]) + _ranges(bs[leaves])).astype(np.int32)
    L.leaf_user = np.concatenate([fmm.cand[b] for b in leaves]).astype(np.int32)
    assert L.leaf_slot.shape[0] == fmm.N and np.unique(L.leaf_user).shape[0] == fmm.N

    # interpolation matrices, one ragged array; matrices a GPU build left in device memory
    # (``gpu_build.DeviceProj``) are gathered there: the device array is the host-built part
    # followed by each result buffer whole, and toff points into it
    L.T, toff = _gather_T(fmm, nb, rk)
    L.toff = toff

    # upward stages: deepest level first; the root level only gathers
    up = []
    for lev in range(fmm.nlevels - 1, fmm.root_level - 1, -1):
        gboxes = [P for P in fmm.gathered_boxes_level(lev)
                  if fmm.child_src[P] is not None and fmm.child_src[P].shape[0] > 0]
        if gboxes:
            gb = np.asarray(gboxes, dtype=np.int64)
            gather_dst = (np.repeat(off[gb], bs[gb]) + _ranges(bs[gb])).astype(np.int32)
            cs = np.concatenate([fmm.child_src[P] for P in gboxes])
            gather_src = (off[cs[:, 0]] + cs[:, 1]).astype(np.int32)
        else:
            gather_dst = gather_src = np.zeros(0, np.int32)
        sboxes = fmm.skel_boxes_level(lev)
        sboxes = sboxes[rk[sboxes] > 0]
        tabs = _level_tables(fmm, sboxes, off, toff)
        items = _box_items(rk[sboxes], ROWS_PER_ITEM)
        up.append(LevelStage(lev, *tabs, gather_dst, gather_src, items))
    L.up = up

    # downward stages: shallowest skeletonized level first
    # slot of child skeleton point (c, j) inside its parent: parents hold their candidates in
    # skeleton-first order, child_src[P][k] = (c, j) for slot off[P] + k; looked up by sorted key
    parents = [P for P in range(nb) if fmm.child_src[P] is not None and fmm.child_src[P].shape[0] > 0 and bs[P] > 0]
    maxj = int(rk.max()) + 1 if rk.size else 1
    if parents:
        cs_all = np.concatenate([fmm.child_src[P] for P in parents]).astype(np.int64)
        pa = np.asarray(parents, dtype=np.int64)
        lens = np.array([fmm.child_src[P].shape[0] for P in parents], dtype=np.int64)
        pos_all = np.repeat(off[pa], lens) + _ranges(lens)
        keys = cs_all[:, 0] * maxj + cs_all[:, 1]
        korder = np.argsort(keys, kind="stable")
        keys_sorted, pos_sorted = keys[korder], pos_all[korder]
    else:
        keys_sorted = pos_sorted = np.zeros(0, np.int64)
    down = []
    for lev in range(fmm.root_level + 1, fmm.nlevels):
        sboxes = fmm.skel_boxes_level(lev)
        sboxes = sboxes[rk[sboxes] > 0]
        r = rk[sboxes].astype(np.int64)
        within = _ranges(r)
        gather_dst = (np.repeat(off[sboxes], r) + within).astype(np.int32)
        q = np.repeat(sboxes.astype(np.int64), r) * maxj + within
        at = np.searchsorted(keys_sorted, q)
        assert np.all(at < keys_sorted.shape[0]) and np.all(keys_sorted[np.minimum(at, max(keys_sorted.shape[0] - 1, 0))] == q)
        gather_src = pos_sorted[at].astype(np.int32)
        tabs = _level_tables(fmm, sboxes, off, toff)
        items = _box_items(bs[sboxes] - rk[sboxes], COLS_PER_ITEM)
        down.append(LevelStage(lev, *tabs, gather_dst, gather_src, items))
    L.down = down

    if near_field_device is not None:
        from .gpu_layout import near_field_tables

        full, split = near_field_tables(fmm, off, bs, rk, L.nslots, max_sources, near_field_device)
        _set_near_field(L, full, split)
    else:
        _near_field(L, fmm, off, bs, rk, max_sources)
    L.max_sources = int(max_sources)
    return L


def _set_near_field(L, full, split):
    for k in NearLists.FIELDS + ("items", "item_cost", "interactions"):
        setattr(L, k, getattr(full, k))
    L.split = split
    L.evaluations = L.interactions if split is None else split[0].interactions + split[1].interactions


def _gather_T(fmm, nb, rk):
    """(T, toff): every box's ``T`` in one array (numpy, or a device tensor when any is on the device)."""
    toff = np.zeros(nb, dtype=np.int64)
    tparts, acc = [], 0
    bufs = {}
    dev = []
    for b in range(nb):
        p = fmm.proj[b]
        if p is None or rk[b] <= 0 or p.size == 0:
            continue
        if hasattr(p, "buf"):  # DeviceProj
            bufs.setdefault(id(p.buf), p.buf)
            dev.append((b, id(p.buf), p.off))
            continue
        toff[b] = acc
        tparts.append(np.ascontiguousarray(p).ravel())
        acc += p.size
    host = np.concatenate(tparts) if tparts else np.zeros(0 if bufs else 1, dtype=fmm.kernel_dtype)
    host = host.astype(fmm.kernel_dtype, copy=False)
    if not bufs:
        return host, toff
    import torch

    cplx = np.dtype(fmm.kernel_dtype).kind == "c"
    w = 2 if cplx else 1
    first = next(iter(bufs.values()))
    parts = [torch.as_tensor(host.view(np.float64) if cplx else host).to(first.device)]
    base, pos = {}, acc
    for k, buf in bufs.items():
        base[k] = pos
        parts.append(buf)
        pos += buf.numel() // w
    for b, k, off in dev:
        toff[b] = base[k] + off
    T = torch.cat(parts)
    return (torch.view_as_complex(T.view(-1, 2)) if cplx else T), toff


def _empty_lists(tpi):
    NL = NearLists()
    for k in NearLists.FIELDS:
        setattr(NL, k, np.zeros(0, dtype=np.int64 if k in ("g_e0", "g_e1") else np.int32))
    NL.items = np.zeros((0, 4), dtype=np.int32)
    NL.item_cost = np.zeros(0, dtype=np.int64)
    NL.interactions = 0
    NL.tpi = tpi
    return NL


def _merge_lists(parts, slot_base, box_base, nslots):
    """Concatenate per-instance near-field lists (objects with the NearLists fields)."""
    M = NearLists()
    g_base = np.cumsum([0] + [P.g_pt.shape[0] for P in parts])[:-1]
    e_base = np.cumsum([0] + [P.e_pt.shape[0] for P in parts])[:-1]
    M.g_pt = np.concatenate([P.g_pt + sb for P, sb in zip(parts, slot_base)]).astype(np.int32)
    M.g_n = np.concatenate([P.g_n for P in parts]).astype(np.int32)
    M.g_kind = np.concatenate([P.g_kind for P in parts]).astype(np.int32)
    M.g_out = (M.g_pt + M.g_kind * nslots).astype(np.int32)
    M.g_box = np.concatenate([P.g_box + bb for P, bb in zip(parts, box_base)]).astype(np.int32)
    M.e_pt = np.concatenate([P.e_pt + sb for P, sb in zip(parts, slot_base)]).astype(np.int32)
    M.e_n = np.concatenate([P.e_n for P in parts]).astype(np.int32)
    M.e_qkind = np.concatenate([P.e_qkind for P in parts]).astype(np.int32)
    M.e_q = (M.e_pt + M.e_qkind * nslots).astype(np.int32)
    M.e_box = np.concatenate([P.e_box + bb for P, bb in zip(parts, box_base)]).astype(np.int32)
    M.e_flags = np.concatenate([P.e_flags for P in parts]).astype(np.int32)
    M.g_e0 = np.concatenate([P.g_e0 + eb for P, eb in zip(parts, e_base)]).astype(np.int64)
    M.g_e1 = np.concatenate([P.g_e1 + eb for P, eb in zip(parts, e_base)]).astype(np.int64)
    items, costs = [], []
    for P, gb, eb in zip(parts, g_base, e_base):
        it = P.items.copy().reshape(-1, 4)
        it[:, 0] += gb
        it[:, 2] += eb
        it[:, 3] += eb
        items.append(it)
        costs.append(P.item_cost)
    items = np.concatenate(items) if items else np.zeros((0, 4), np.int32)
    costs = np.concatenate(costs) if costs else np.zeros(0, np.int64)
    order = np.argsort(-costs, kind="stable")
    M.items = np.ascontiguousarray(items[order]).astype(np.int32)
    M.item_cost = costs[order]
    M.interactions = int(sum(P.interactions for P in parts))
    return M


def merge_layouts(layouts):
    """One layout for several independent instances (same operator and dimension).

    Slots, boxes, points, T and work lists are concatenated with offsets; the level
    stages of all instances that sit on the same tree level are merged into one stage,
    so one upward and one downward launch per level serve every instance (each
    instance still runs its levels deepest first upward and shallowest first downward).
    """
    if len(layouts) == 1:
        return layouts[0]
    M = BoxLayout()
    M.ndim = layouts[0].ndim
    M.N = sum(L.N for L in layouts)
    M.nboxes = sum(L.nboxes for L in layouts)
    M.nslots = sum(L.nslots for L in layouts)
    slot_base = np.cumsum([0] + [L.nslots for L in layouts])[:-1]
    box_base = np.cumsum([0] + [L.nboxes for L in layouts])[:-1]
    pt_base = np.cumsum([0] + [L.N for L in layouts])[:-1]
    t_base = np.cumsum([0] + [L.T.shape[0] for L in layouts])[:-1]
    M.instance_points = np.array([L.N for L in layouts], dtype=np.int64)
    M.instance_point_base = pt_base.astype(np.int64)
    M.off = np.concatenate([np.where(L.off >= 0, L.off + sb, -1) for L, sb in zip(layouts, slot_base)])
    M.slot_point = np.concatenate([L.slot_point + pb for L, pb in zip(layouts, pt_base)])
    M.slot_box = np.concatenate([L.slot_box + bb for L, bb in zip(layouts, box_base)])
    M.points = np.concatenate([L.points for L in layouts], axis=1)
    M.normals = None if layouts[0].normals is None else np.concatenate([L.normals for L in layouts], axis=1)
    M.centers = np.concatenate([L.centers for L in layouts])
    M.leaf_slot = np.concatenate([L.leaf_slot + sb for L, sb in zip(layouts, slot_base)]).astype(np.int32)
    M.leaf_user = np.concatenate([L.leaf_user + pb for L, pb in zip(layouts, pt_base)]).astype(np.int32)
    if any(not isinstance(L.T, np.ndarray) for L in layouts):  # some T on the device
        import torch

        dev = next(L.T.device for L in layouts if not isinstance(L.T, np.ndarray))
        M.T = torch.cat([torch.as_tensor(L.T).to(dev) for L in layouts])
    else:
        M.T = np.concatenate([L.T for L in layouts])
    M.toff = np.concatenate([L.toff + tb for L, tb in zip(layouts, t_base)])
    M.root_level = min(L.root_level for L in layouts)

    def merge_stages(stage_lists, descending):
        by_level = {}
        for i, stages in enumerate(stage_lists):
            for S in stages:
                by_level.setdefault(S.level, []).append((i, S))
        out = []
        for lev in sorted(by_level, reverse=descending):
            parts = by_level[lev]
            boxes, off, bsz, rank, toff, gd, gs, items = [], [], [], [], [], [], [], []
            nbox = 0
            for i, S in parts:
                boxes.append(S.boxes + box_base[i])
                off.append(S.box_off + slot_base[i])
                bsz.append(S.box_bs)
                rank.append(S.box_rank)
                toff.append(S.box_toff + t_base[i])
                gd.append(S.gather_dst + slot_base[i])
                gs.append(S.gather_src + slot_base[i])
                it = S.items.copy()
                it[:, 0] += nbox
                items.append(it)
                nbox += S.box_off.shape[0]
            out.append(LevelStage(lev, np.concatenate(boxes), np.concatenate(off).astype(np.int32),
                                  np.concatenate(bsz).astype(np.int32), np.concatenate(rank).astype(np.int32),
                                  np.concatenate(toff).astype(np.int64), np.concatenate(gd).astype(np.int32),
                                  np.concatenate(gs).astype(np.int32),
                                  np.concatenate(items).astype(np.int32).reshape(-1, 2)))
        return out

    M.up = merge_stages([L.up for L in layouts], True)
    M.down = merge_stages([L.down for L in layouts], False)

    full = _merge_lists(layouts, slot_base, box_base, M.nslots)
    for k in NearLists.FIELDS + ("items", "item_cost"):
        setattr(M, k, getattr(full, k))
    M.split = None
    if any(L.split is not None for L in layouts):
        ones = [L.split[0] if L.split is not None else L for L in layouts]
        syms = [L.split[1] if L.split is not None else _empty_lists(SYM_TARGETS_PER_ITEM) for L in layouts]
        M.split = (_merge_lists(ones, slot_base, box_base, M.nslots),
                   _merge_lists(syms, slot_base, box_base, M.nslots))
        M.split[0].tpi, M.split[1].tpi = TARGETS_PER_ITEM, SYM_TARGETS_PER_ITEM
    M.evaluations = int(sum(L.evaluations for L in layouts))
    M.interactions = int(sum(L.interactions for L in layouts))
    M.max_sources = int(layouts[0].max_sources)
    return M


def make_items(g_n, g_e0, g_e1, e_n, max_sources, targets_per_item=TARGETS_PER_ITEM):
    """Near-field work items {group, first target, first entry, end entry}, heaviest first.

    Each item covers up to ``targets_per_item`` targets of one group and a run of that
    group's consecutive source entries: the entries whose first source point falls in the
    same window of ``max_sources`` points of the group's source list (a run holds at most
    ``max_sources`` plus one entry's points, and at least one entry).
    """
    g_n = np.asarray(g_n, dtype=np.int64)
    g_e0 = np.asarray(g_e0, dtype=np.int64)
    g_e1 = np.asarray(g_e1, dtype=np.int64)
    e_n = np.asarray(e_n, dtype=np.int64)
    ne = g_e1 - g_e0
    if g_n.shape[0] == 0 or ne.sum() == 0:
        return np.zeros((0, 4), dtype=np.int32), np.zeros(0, dtype=np.int64)
    cums = np.concatenate(([0], np.cumsum(e_n)))
    gid = np.repeat(np.arange(g_n.shape[0], dtype=np.int64), ne)
    ent = np.repeat(g_e0, ne) + _ranges(ne)
    loc = cums[ent] - cums[g_e0[gid]]
    bucket = loc // int(max_sources)
    new_run = np.ones(ent.shape[0], dtype=bool)
    new_run[1:] = (gid[1:] != gid[:-1]) | (bucket[1:] != bucket[:-1])
    ra = ent[new_run]
    rg = gid[new_run]
    rb = np.concatenate((ra[1:], [0]))
    last_of_group = np.concatenate((rg[1:] != rg[:-1], [True]))
    rb[last_of_group] = g_e1[rg[last_of_group]]
    rsrc = cums[rb] - cums[ra]
    nblk = (g_n[rg] + targets_per_item - 1) // targets_per_item
    item_run = np.repeat(np.arange(ra.shape[0]), nblk)
    t0 = _ranges(nblk) * targets_per_item
    ig = rg[item_run]
    items = np.stack([ig, t0, ra[item_run], rb[item_run]], axis=1).astype(np.int32)
    costs = np.minimum(targets_per_item, g_n[ig] - t0) * rsrc[item_run]
    order = np.argsort(-costs, kind="stable")
    return np.ascontiguousarray(items[order]), costs[order].astype(np.int64)


ITEM_ORDERS = ("heavy", "xcd")
# XCDs of the device: workgroup i of a launch runs on XCD i mod 8, each XCD with an L2 of its own
N_XCD = 8


def order_items(items, costs, mode="heavy", nxcd=N_XCD, chunks_per_xcd=4):
    """Launch order of near-field work items (workgroup i takes item i).

    ``"heavy"``: heaviest first (the stable cost order of ``make_items``), which shortens the
    tail of the launch. ``"xcd"``: items in table order (group by group, i.e. box by box in
    Morton order within a level, target blocks and source runs in turn) cut into
    ``nxcd * chunks_per_xcd`` runs of equal cost; run c goes to XCD c mod nxcd and workgroup
    ``x + nxcd p`` takes item p of XCD x's runs, so the items one XCD executes together are
    spatial neighbours that read the same source sets through its L2. XCDs with fewer items are
    padded with empty items ``{0, 0, 0, 0}`` (no source entries), which the kernels skip.
    Returns (items, costs) as new arrays.
    """
    items = np.asarray(items, dtype=np.int32).reshape(-1, 4)
    costs = np.asarray(costs, dtype=np.int64)
    if mode not in ITEM_ORDERS:
        raise ValueError(f"item order must be one of {ITEM_ORDERS}, got {mode!r}")
    n = items.shape[0]
    if mode == "heavy" or n == 0:
        o = np.argsort(-costs, kind="stable")
        return np.ascontiguousarray(items[o]), costs[o]
    o = np.lexsort((items[:, 3], items[:, 2], items[:, 1], items[:, 0]))
    items, costs = items[o], costs[o]
    nrun = nxcd * chunks_per_xcd
    cum = np.cumsum(costs)
    total = max(int(cum[-1]), 1)
    cut = np.searchsorted(cum, total * np.arange(1, nrun, dtype=np.float64) / nrun, side="left")
    bounds = np.concatenate(([0], np.clip(cut, 0, n), [n]))
    seqs = [np.concatenate([np.arange(bounds[c], bounds[c + 1]) for c in range(x, nrun, nxcd)])
            for x in range(nxcd)]
    L = max(s.shape[0] for s in seqs)
    pos = np.full((L, nxcd), -1, dtype=np.int64)
    for x, s in enumerate(seqs):
        pos[:s.shape[0], x] = s
    pos = pos.reshape(-1)
    # drop trailing empty slots (nothing after them needs its XCD position)
    last = np.nonzero(pos >= 0)[0][-1]
    pos = pos[:last + 1]
    out = np.zeros((pos.shape[0], 4), dtype=np.int32)
    oc = np.zeros(pos.shape[0], dtype=np.int64)
    real = pos >= 0
    out[real] = items[pos[real]]
    oc[real] = costs[pos[real]]
    return out, oc


def resplit(L, max_sources):
    """Rebuild a layout's near-field work items for another source-run length (in place)."""
    L.items, L.item_cost = make_items(L.g_n, L.g_e0, L.g_e1, L.e_n, max_sources)
    if L.split is not None:
        for part in L.split:
            part.items, part.item_cost = make_items(part.g_n, part.g_e0, part.g_e1, part.e_n, max_sources, part.tpi)
    L.max_sources = int(max_sources)
    return L


class NearLists:
    """Near-field groups (target sets), entries (source sets) and work items."""

    FIELDS = ("g_pt", "g_n", "g_out", "g_box", "g_kind", "g_e0", "g_e1", "e_pt", "e_n", "e_q", "e_qkind", "e_box",
              "e_flags")


# kernels whose pairs of lists (rows U -> V and V -> U) are evaluated once per point pair by the
# symmetric near-field kernel: single layers, K(x, y) = K(y, x), and the double layers,
# K(x, y) = g(|x - y|) (x - y) . n with n the normal at y (double layer) or x (adjoint), whose
# two directions share g and differ in the projection
SYMMETRIC_KERNEL_IDS = (0, 1, 2, 3, 4, 5, 6, 7)
SYM_TARGETS_PER_ITEM = 512


def _lists_from_rows(R, off, bs, rk, nslots, max_sources, tpi):
    """Groups / entries / items from rows (tbox, tkind, sbox, skind, qkind, neg) with nonzero sets."""
    NL = NearLists()
    NL.tpi = int(tpi)
    order = np.lexsort((R[:, 2], R[:, 1], R[:, 0]))
    R = R[order]
    scount = np.where(R[:, 3] == 0, bs[R[:, 2]], rk[R[:, 2]])
    self_pair = (R[:, 0] == R[:, 2]) & (R[:, 1] == R[:, 3])
    NL.e_pt = off[R[:, 2]].astype(np.int32)
    NL.e_n = scount.astype(np.int32)
    NL.e_qkind = R[:, 4].astype(np.int32)
    NL.e_q = (off[R[:, 2]] + R[:, 4] * nslots).astype(np.int32)
    NL.e_flags = (R[:, 5] * 1 + self_pair * 2).astype(np.int32)
    NL.e_box = R[:, 2].astype(np.int32)
    key = R[:, 0] * 2 + R[:, 1]
    gkeys, gstart, gcount = np.unique(key, return_index=True, return_counts=True)
    g_box = gkeys // 2
    g_kind = gkeys % 2
    NL.g_pt = off[g_box].astype(np.int32)
    NL.g_n = np.where(g_kind == 0, bs[g_box], rk[g_box]).astype(np.int32)
    NL.g_out = (off[g_box] + g_kind * nslots).astype(np.int32)
    NL.g_box = g_box.astype(np.int32)
    NL.g_kind = g_kind.astype(np.int32)
    NL.g_e0 = gstart.astype(np.int64)
    NL.g_e1 = (gstart + gcount).astype(np.int64)
    NL.items, NL.item_cost = make_items(NL.g_n, NL.g_e0, NL.g_e1, NL.e_n, max_sources, tpi)
    tn = NL.g_n[np.searchsorted(gkeys, key)].astype(np.int64) if key.shape[0] else np.zeros(0, np.int64)
    NL.interactions = int(np.sum(tn * NL.e_n.astype(np.int64)))
    return NL


# the symmetric kernel takes a work item's targets in rows of this many lanes
SYM_LANES = 64


def pair_representative(a, b, sizes):
    """Which row of a pair is its representative (array of bools over the rows).

    Without ``sizes``: the row whose target set orders first (``a < b``). With ``sizes`` (per row,
    the target and source set sizes): the row whose target set leaves fewer idle lanes times
    sources in the symmetric kernel's rows of ``SYM_LANES`` targets, ties by order. The rule is
    antisymmetric, so exactly one row of every pair is chosen; host and device builds share it.
    """
    if sizes is None:
        return a < b
    nt, ns = sizes
    wt = ((nt + SYM_LANES - 1) // SYM_LANES * SYM_LANES - nt) * ns  # this row as targets
    ws = ((ns + SYM_LANES - 1) // SYM_LANES * SYM_LANES - ns) * nt  # the mirror row as targets
    return (wt < ws) | ((wt == ws) & (a < b))


def pair_rows(R, nboxes, sizes=None):
    """Split rows into (one-sided rows, pair representatives).

    Row (U, a, V, b, q, neg) pairs with (V, b, U, a, q', neg) for U != V when both are unique
    and each charge kind equals its source set's kind (a set's charges and potentials then share
    slots). The representative follows ``pair_representative``.
    """
    if R.shape[0] == 0:
        return R, R
    a = R[:, 0] * 2 + R[:, 1]
    b = R[:, 2] * 2 + R[:, 3]
    nb2 = 2 * int(nboxes)
    key = (a * nb2 + b) * 2 + R[:, 5]
    pkey = (b * nb2 + a) * 2 + R[:, 5]
    uk, cnt = np.unique(key, return_counts=True)
    single = np.isin(key, uk[cnt == 1])
    ok = single & (R[:, 0] != R[:, 2]) & (R[:, 4] == R[:, 3])
    okkeys = key[ok]
    paired = ok & np.isin(pkey, okkeys)
    rep = paired & pair_representative(a, b, sizes)
    return R[~paired], R[rep]


_GROUP_FIELDS = ("g_pt", "g_n", "g_out", "g_box", "g_kind")
_ENTRY_FIELDS = ("e_pt", "e_n", "e_q", "e_qkind", "e_box", "e_flags")


def _assemble(groups, counts, entries, max_sources, tpi):
    """NearLists from group columns, each group's entry count and the entry columns (in group order)."""
    NL = NearLists()
    NL.tpi = int(tpi)
    for k in _GROUP_FIELDS:
        setattr(NL, k, np.asarray(groups[k], dtype=np.int32))
    for k in _ENTRY_FIELDS:
        setattr(NL, k, np.asarray(entries[k], dtype=np.int32))
    counts = np.asarray(counts, dtype=np.int64)
    NL.g_e1 = np.cumsum(counts).astype(np.int64)
    NL.g_e0 = NL.g_e1 - counts
    NL.items, NL.item_cost = make_items(NL.g_n, NL.g_e0, NL.g_e1, NL.e_n, max_sources, tpi)
    tn = np.repeat(NL.g_n.astype(np.int64), counts)
    NL.interactions = int(np.sum(tn * NL.e_n.astype(np.int64)))
    return NL


def split_self_blocks(one, sym, max_sources, block):
    """Symmetric evaluation inside the self segments: (one, sym) with every self segment (a
    target set against its own points, one-sided with SRC_SELF) of more than ``block`` points cut
    into blocks of ``block`` targets. The pairs of two different blocks move to the symmetric
    table -- block i as targets against the points after it, one group per block, evaluated once
    for both directions -- and each diagonal block stays one-sided as a group of its own (self
    flag kept, so i == j is skipped). A segment qualifies when its charge slots are its potential
    slots. Same interactions: n^2 = sum of diagonal blocks + 2 x the pairs of different blocks.
    ``block`` 0 returns (one, sym) unchanged; multiples of ``SYM_LANES`` leave no idle lanes in
    the symmetric kernel's rows."""
    block = int(block)
    if block <= 0:
        return one, sym
    col = lambda P, k: np.asarray(getattr(P, k), dtype=np.int64)
    g = {k: col(one, k) for k in _GROUP_FIELDS}
    e = {k: col(one, k) for k in _ENTRY_FIELDS}
    ne = col(one, "g_e1") - col(one, "g_e0")
    eg = np.repeat(np.arange(ne.shape[0], dtype=np.int64), ne)  # group of every entry
    sel = (((e["e_flags"] & SRC_SELF) != 0) & (e["e_n"] > block) & (e["e_pt"] == g["g_pt"][eg])
           & (e["e_n"] == g["g_n"][eg]) & (e["e_q"] == g["g_out"][eg]))
    if not sel.any():
        return one, sym
    s = np.nonzero(sel)[0]
    n = e["e_n"][s]
    # diagonal blocks: one group and one self entry each
    nb = (n + block - 1) // block
    ds = np.repeat(s, nb)
    d0 = _ranges(nb) * block
    dn = np.minimum(block, np.repeat(n, nb) - d0)
    # pairs of different blocks: block i (full) against the points after it
    ps = np.repeat(s, nb - 1)
    p0 = _ranges(nb - 1) * block
    pn = np.repeat(n, nb - 1) - p0 - block

    def sub_groups(ent, start, cnt):
        gi = eg[ent]
        return {"g_pt": g["g_pt"][gi] + start, "g_n": cnt, "g_out": g["g_out"][gi] + start,
                "g_box": g["g_box"][gi], "g_kind": g["g_kind"][gi]}

    def sub_entries(ent, start, cnt, flags):
        return {"e_pt": e["e_pt"][ent] + start, "e_n": cnt, "e_q": e["e_q"][ent] + start,
                "e_qkind": e["e_qkind"][ent], "e_box": e["e_box"][ent], "e_flags": flags}

    keep = ~sel
    kept_ne = np.bincount(eg[keep], minlength=ne.shape[0])
    live = kept_ne > 0
    dg = sub_groups(ds, d0, dn)
    de = sub_entries(ds, d0, dn, e["e_flags"][ds])
    one2 = _assemble({k: np.concatenate((g[k][live], dg[k])) for k in _GROUP_FIELDS},
                     np.concatenate((kept_ne[live], np.ones(ds.shape[0], np.int64))),
                     {k: np.concatenate((e[k][keep], de[k])) for k in _ENTRY_FIELDS}, max_sources, one.tpi)
    sg = sub_groups(ps, p0, np.full(ps.shape[0], block, np.int64))
    se = sub_entries(ps, p0 + block, pn, e["e_flags"][ps] & ~SRC_SELF)
    sym_ne = col(sym, "g_e1") - col(sym, "g_e0")
    sym2 = _assemble({k: np.concatenate((col(sym, k), sg[k])) for k in _GROUP_FIELDS},
                     np.concatenate((sym_ne, np.ones(ps.shape[0], np.int64))),
                     {k: np.concatenate((col(sym, k), se[k])) for k in _ENTRY_FIELDS}, max_sources, sym.tpi)
    assert one2.interactions + 2 * sym2.interactions == one.interactions + 2 * sym.interactions
    return one2, sym2


def _near_field(L, fmm, off, bs, rk, max_sources):
    nslots = L.nslots
    # (target box, target kind) -> list of (source box, source kind, charge kind, flags); kind 0 = all, 1 = skel
    add_lists = [fmm.u_from_qskel_list, fmm.u_from_q_list, fmm.tree_uskel_from_qskel_list,
                 fmm.uskel_from_q_list, fmm.leaf_uskel_from_qskel_list]
    rows = []  # columns: tbox, tkind, sbox, skind, qkind, neg
    for lst in add_lists:
        if lst.shape[0]:
            n = lst.shape[0]
            rows.append(np.stack([lst[:, 0], np.zeros(n, np.int64), lst[:, 1], np.zeros(n, np.int64),
                                  np.zeros(n, np.int64), np.zeros(n, np.int64)], axis=1))
    lst = fmm.uskel_from_qskel_list
    if lst.shape[0]:
        n = lst.shape[0]
        rows.append(np.stack([lst[:, 0], np.ones(n, np.int64), lst[:, 1], np.ones(n, np.int64),
                              np.ones(n, np.int64), np.ones(n, np.int64)], axis=1))
    lst = fmm.uskel_from_q_list
    if lst.shape[0]:
        n = lst.shape[0]
        rows.append(np.stack([lst[:, 0], np.ones(n, np.int64), lst[:, 1], np.zeros(n, np.int64),
                              np.zeros(n, np.int64), np.ones(n, np.int64)], axis=1))
    lst = fmm.u_from_qskel_list
    if lst.shape[0]:
        n = lst.shape[0]
        rows.append(np.stack([lst[:, 0], np.zeros(n, np.int64), lst[:, 1], np.ones(n, np.int64),
                              np.ones(n, np.int64), np.ones(n, np.int64)], axis=1))
    R = np.concatenate(rows) if rows else np.zeros((0, 6), np.int64)

    tcount = np.where(R[:, 1] == 0, bs[R[:, 0]], rk[R[:, 0]])
    scount = np.where(R[:, 3] == 0, bs[R[:, 2]], rk[R[:, 2]])
    keep = (tcount > 0) & (scount > 0)
    R, tcount, scount = R[keep], tcount[keep].astype(np.int64), scount[keep].astype(np.int64)

    full = _lists_from_rows(R, off, bs, rk, nslots, max_sources, TARGETS_PER_ITEM)
    for k in NearLists.FIELDS + ("items", "item_cost", "interactions"):
        setattr(L, k, getattr(full, k))
    L.evaluations = L.interactions
    L.split = None
    if fmm.operator.kernel_id in SYMMETRIC_KERNEL_IDS:
        one_rows, sym_rows = pair_rows(R, fmm.nboxes, sizes=(tcount, scount))
        if sym_rows.shape[0]:
            one = _lists_from_rows(one_rows, off, bs, rk, nslots, max_sources, TARGETS_PER_ITEM)
            sym = _lists_from_rows(sym_rows, off, bs, rk, nslots, max_sources, SYM_TARGETS_PER_ITEM)
            assert one.interactions + 2 * sym.interactions == L.interactions
            L.split = (one, sym)
            L.evaluations = one.interactions + sym.interactions


def _fail(what):
    raise ValueError(f"inconsistent box layout: {what}")


def _within(name, a, lo, hi):
    """Every entry of ``a`` in ``[lo, hi)`` (``hi`` may be an array of the same length)."""
    a = np.asarray(a, dtype=np.int64)
    if a.size and (a.min() < lo or np.any(a >= hi)):
        _fail(f"{name} outside [{lo}, {hi if np.isscalar(hi) else 'limit'})")


def _lengths(name, n, *arrays):
    for a in arrays:
        if a is None or np.asarray(a).shape[0] != n:
            _fail(f"{name}: tables of different lengths")


def _check_lists(name, P, nslots, nboxes):
    ng, ne = int(np.asarray(P.g_pt).shape[0]), int(np.asarray(P.e_pt).shape[0])
    _lengths(name + " groups", ng, P.g_pt, P.g_n, P.g_out, P.g_box, P.g_e0, P.g_e1)
    _lengths(name + " entries", ne, P.e_pt, P.e_n, P.e_q, P.e_box, P.e_flags)
    g_pt, g_n, g_out = (np.asarray(a, dtype=np.int64) for a in (P.g_pt, P.g_n, P.g_out))
    e_pt, e_n, e_q = (np.asarray(a, dtype=np.int64) for a in (P.e_pt, P.e_n, P.e_q))
    g_e0, g_e1 = np.asarray(P.g_e0, dtype=np.int64), np.asarray(P.g_e1, dtype=np.int64)
    _within(name + ".g_n", g_n, 0, nslots + 1)
    _within(name + ".g_pt", g_pt, 0, nslots - g_n + 1)
    _within(name + ".g_out", g_out, 0, 2 * nslots - g_n + 1)
    _within(name + ".g_box", P.g_box, 0, nboxes)
    _within(name + ".e_n", e_n, 0, nslots + 1)
    _within(name + ".e_pt", e_pt, 0, nslots - e_n + 1)
    _within(name + ".e_q", e_q, 0, 2 * nslots - e_n + 1)
    _within(name + ".e_box", P.e_box, 0, nboxes)
    _within(name + ".e_flags", P.e_flags, 0, SRC_NEG + SRC_SELF + 1)
    _within(name + ".g_e0", g_e0, 0, ne + 1)
    if ng and (np.any(g_e1 < g_e0) or np.any(g_e1 > ne)):
        _fail(f"{name}: group entry ranges outside the entry table")
    it = np.asarray(P.items, dtype=np.int64).reshape(-1, 4)
    if it.shape[0]:
        _within(name + " item group", it[:, 0], 0, ng)
        g = it[:, 0]
        if (np.any(it[:, 1] < 0) or np.any(it[:, 1] >= np.maximum(g_n[g], 1))
                or np.any(it[:, 2] < g_e0[g]) or np.any(it[:, 3] > g_e1[g]) or np.any(it[:, 2] > it[:, 3])):
            _fail(f"{name}: a work item leaves its group's targets or entries")


def check_layout(L):
    """Raise ``ValueError`` unless every index of layout ``L`` stays inside the arrays it indexes.

    The device kernels trust these tables (slot ranges, charge/potential offsets, T offsets,
    gather indices, work items); a layout read from a file is checked here before upload."""
    ndim, N, nboxes, nslots = int(L.ndim), int(L.N), int(L.nboxes), int(L.nslots)
    if ndim not in (2, 3) or min(N, nboxes, nslots) < 0:
        _fail("bad dimension or sizes")
    pts = np.asarray(L.points)
    if pts.shape != (ndim, nslots):
        _fail(f"points shape {pts.shape}, expected {(ndim, nslots)}")
    if L.normals is not None and np.asarray(L.normals).shape != (3, nslots):
        _fail("normals shape")
    if np.asarray(L.centers).shape != (nboxes, 3):
        _fail("centers shape")
    _lengths("slot tables", nslots, L.slot_box, L.slot_point)
    _within("slot_box", L.slot_box, 0, nboxes)
    _within("slot_point", L.slot_point, 0, N)
    _lengths("box offsets", nboxes, L.off, L.toff)
    _within("off", L.off, -1, nslots)
    _lengths("leaf tables", N, L.leaf_slot, L.leaf_user)
    _within("leaf_slot", L.leaf_slot, 0, nslots)
    _within("leaf_user", L.leaf_user, 0, N)
    ntt = int(L.T.shape[0])
    _within("toff", L.toff, 0, ntt + 1)
    for direction in ("up", "down"):
        for S in getattr(L, direction):
            name = f"{direction} stage (level {S.level})"
            nb = int(np.asarray(S.box_off).shape[0])
            _lengths(name, nb, S.box_bs, S.box_rank, S.box_toff)
            off, bs, rk, toff = (np.asarray(a, dtype=np.int64) for a in (S.box_off, S.box_bs, S.box_rank,
                                                                         S.box_toff))
            if nb and (np.any(rk < 0) or np.any(rk > bs) or np.any(off < 0) or np.any(off + bs > nslots)
                       or np.any(toff < 0) or np.any(toff + rk * (bs - rk) > ntt)):
                _fail(f"{name}: a box's slots or T block leave their arrays")
            _lengths(name + " gathers", int(np.asarray(S.gather_dst).shape[0]), S.gather_src)
            _within(name + " gather_dst", S.gather_dst, 0, nslots)
            _within(name + " gather_src", S.gather_src, 0, nslots)
            it = np.asarray(S.items, dtype=np.int64).reshape(-1, 2)
            _within(name + " item box", it[:, 0], 0, nb)
            _within(name + " item start", it[:, 1], 0, nslots + 1)
    _check_lists("near field", L, nslots, nboxes)
    for i, part in enumerate(getattr(L, "split", None) or ()):
        _check_lists(f"split part {i}", part, nslots, nboxes)
    return L
