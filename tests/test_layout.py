"""Box layout and build invariants (CPU): every map the HIP kernels consume is checked here."""

import numpy as np
import pytest

from conftest import annulus_points, make_build, rel_l2, sphere_points
from layout_emulator import emulate_apply
from oracle.batched_apply import BatchedApplyCPU
from skelfmm_amd import operators as ops
from skelfmm_amd.layout import N_XCD, SRC_NEG, SRC_SELF, SYMMETRIC_KERNEL_IDS, build_layout, order_items


def _builds():
    rng = np.random.default_rng(0)
    yield "lap2d", make_build(rng.random((3000, 2)), ops.LAPLACE_2D, leaf_size=50), 0.0
    yield "annulus", make_build(annulus_points(4000, 3), ops.LAPLACE_2D, tol=1e-5, leaf_size=40), 0.0
    c = np.vstack([rng.random((2500, 2)) * 0.05, rng.random((1500, 2))])
    yield "clustered2d", make_build(c, ops.LAPLACE_2D, leaf_size=30), 0.0
    yield "lap3d", make_build(rng.random((8000, 3)), ops.LAPLACE_3D, leaf_size=100), 0.0
    v = sphere_points(3000, 1)
    yield "ldl", make_build(v, ops.LAPLACE_DOUBLE_LAYER_3D, leaf_size=60, geometry=v), 0.0


BUILDS = list(_builds())


@pytest.mark.parametrize("case", BUILDS, ids=lambda c: c[0])
def test_build_structure(case):
    """test_fmm.py:111-144: root level, zero ranks at root boxes, list levels."""
    _, f, _ = case
    t = f.tree
    assert f.root_level <= 2 or f.nlevels - 1 < 2
    for b in f.root_boxes:
        assert f.rank_list[b] == 0
        assert t.get_box_level(b) <= f.root_level
    for b in range(f.nboxes):
        if t.get_box_level(b) > f.root_level:
            assert f.rank_list[b] > 0
            assert f.proj[b].shape == (f.rank_list[b], f.bs_list[b] - f.rank_list[b])
    for U, Q in f.uskel_from_qskel_list:
        assert t.get_box_level(U) == t.get_box_level(Q) > f.root_level
    for U, C in f.uskel_from_q_list:
        assert t.get_box_level(U) > f.root_level and t.is_leaf(C) and t.get_box_level(C) < t.get_box_level(U)
    assert np.array_equal(f.u_from_qskel_list, f.uskel_from_q_list[:, ::-1])
    # every point pair is covered: leaf candidates partition the points
    allp = np.concatenate([f.cand[b] for b in t.get_leaves()])
    assert np.array_equal(np.sort(allp), np.arange(f.N))


@pytest.mark.parametrize("case", BUILDS, ids=lambda c: c[0])
def test_layout_maps(case):
    _, f, _ = case
    L = build_layout(f)
    # slots hold the candidates, skeleton first
    for b in range(f.nboxes):
        if f.bs_list[b]:
            assert np.array_equal(L.slot_point[L.off[b]:L.off[b] + f.bs_list[b]], f.cand[b])
    assert np.array_equal(np.sort(L.leaf_user), np.arange(f.N))
    assert np.array_equal(L.slot_point[L.leaf_slot], L.leaf_user)
    for S in L.up:
        # a parent's candidate slot holds the same point as the child skeleton slot it is gathered from
        assert np.array_equal(L.slot_point[S.gather_dst], L.slot_point[S.gather_src])
        assert np.all(S.box_rank > 0)
    for S in L.down:
        assert np.array_equal(L.slot_point[S.gather_dst], L.slot_point[S.gather_src])
        assert np.all(L.slot_box[S.gather_src] != L.slot_box[S.gather_dst])
    assert [S.level for S in L.up] == sorted([S.level for S in L.up], reverse=True)
    assert [S.level for S in L.down] == sorted([S.level for S in L.down])


@pytest.mark.parametrize("case", BUILDS, ids=lambda c: c[0])
def test_work_items_cover_every_interaction_once(case):
    _, f, _ = case
    L = build_layout(f, max_sources=700)
    cover = {}
    for g, t0, sb, se in L.items:
        assert 0 <= t0 < L.g_n[g] and t0 % 64 == 0 and sb < se
        for s in range(sb, se):
            cover[(g, t0, s)] = cover.get((g, t0, s), 0) + 1
    ngroups = L.g_pt.shape[0]
    starts = {}
    for (g, t0, s) in cover:
        starts.setdefault(g, set()).add(s)
    for g in range(ngroups):
        chunks = list(range(0, int(L.g_n[g]), 64))
        ents = sorted(starts[g])
        assert ents == list(range(ents[0], ents[-1] + 1))
        for t0 in chunks:
            for s in ents:
                assert cover[(g, t0, s)] == 1
    # self entries only where a group meets its own slots, and the counts are right
    for g in range(ngroups):
        for s in starts[g]:
            if L.e_flags[s] & SRC_SELF:
                assert L.e_pt[s] == L.g_pt[g] and L.e_n[s] == L.g_n[g]
    neg = (L.e_flags & SRC_NEG) != 0
    nsub = f.uskel_from_qskel_list.shape[0] + 2 * f.uskel_from_q_list.shape[0]
    assert neg.sum() <= nsub
    assert np.all(L.item_cost[:-1] >= L.item_cost[1:])


@pytest.mark.parametrize("case", BUILDS, ids=lambda c: c[0])
def test_layout_emulation_matches_oracle(case):
    """The kernels' work lists, run in numpy, reproduce the oracle's padded apply to 1e-13."""
    _, f, k = case
    L = build_layout(f, max_sources=900)
    q = np.random.RandomState(4).randn(f.N)
    y = emulate_apply(L, f.operator.name, k, q)
    ref = BatchedApplyCPU(f, bytes_budget=1e8).matvec(q).numpy()
    assert rel_l2(y, ref) < 1e-13


def test_interaction_count_matches_lists():
    f = BUILDS[0][1]
    L = build_layout(f)
    bs, rk = f.bs_list, f.rank_list
    n = 0
    for lst in (f.u_from_qskel_list, f.u_from_q_list, f.tree_uskel_from_qskel_list, f.uskel_from_q_list,
                f.leaf_uskel_from_qskel_list):
        n += int(np.sum(bs[lst[:, 0]] * bs[lst[:, 1]])) if lst.shape[0] else 0
    lst = f.uskel_from_qskel_list
    n += int(np.sum(rk[lst[:, 0]] * rk[lst[:, 1]]))
    lst = f.uskel_from_q_list
    n += 2 * int(np.sum(rk[lst[:, 0]] * bs[lst[:, 1]])) if lst.shape[0] else 0
    assert L.interactions == n


def test_merged_layout_emulates_every_instance():
    """Three independent instances in one layout reproduce their separate oracle applies."""
    from skelfmm_amd.layout import merge_layouts

    rng = np.random.default_rng(8)
    builds = [make_build(annulus_points(n, s), ops.LAPLACE_2D, tol=1e-5, leaf_size=40)
              for n, s in ((1500, 1), (4000, 2), (900, 3))]
    L = merge_layouts([build_layout(b, max_sources=600) for b in builds])
    assert L.N == sum(b.N for b in builds)
    qs = [rng.normal(size=b.N) for b in builds]
    y = emulate_apply(L, "laplace_2d", 0.0, np.concatenate(qs))
    off = 0
    for b, q in zip(builds, qs):
        ref = BatchedApplyCPU(b, bytes_budget=1e8).matvec(q).numpy()
        assert rel_l2(y[off:off + b.N], ref) < 1e-13
        off += b.N
    levels = [S.level for S in L.up]
    assert levels == sorted(set(levels), reverse=True)


SYMMETRIC_BUILDS = [c for c in BUILDS if c[1].operator.kernel_id in SYMMETRIC_KERNEL_IDS] + [
    ("helm3d", make_build(np.random.default_rng(2).random((2500, 3)), ops.HELMHOLTZ_3D, kappa=3.0, leaf_size=40), 3.0),
    ("hadl", make_build(sphere_points(2000, 3), ops.HELMHOLTZ_ADJOINT_DOUBLE_LAYER_3D, kappa=3.0, leaf_size=50,
                        geometry=sphere_points(2000, 3)), 3.0)]


@pytest.mark.parametrize("case", SYMMETRIC_BUILDS, ids=lambda c: c[0])
def test_symmetric_split_emulation_matches_one_sided(case):
    """Pairs evaluated once in both directions plus the one-sided rest give the same apply."""
    _, f, k = case
    L = build_layout(f, max_sources=900)
    assert L.split is not None
    one, sym = L.split
    assert one.interactions + 2 * sym.interactions == L.interactions
    assert L.evaluations == one.interactions + sym.interactions < L.interactions
    assert not np.any(sym.e_flags & 2), "self segments must stay one-sided"
    assert np.all(sym.e_qkind == np.where(sym.e_q >= L.nslots, 1, 0))
    q = np.random.RandomState(4).randn(f.N)
    if f.operator.is_complex:
        q = q + 1j * np.random.RandomState(5).randn(f.N)
    y0 = emulate_apply(L, f.operator.name, k, q)
    y1 = emulate_apply(L, f.operator.name, k, q, split=True)
    assert rel_l2(y1, y0) < 1e-13


@pytest.mark.parametrize("block", [16, 24])
@pytest.mark.parametrize("case", [c for c in SYMMETRIC_BUILDS if c[0] in ("helm3d", "hadl")] + SYMMETRIC_BUILDS[:1],
                         ids=lambda c: c[0])
def test_self_blocks_emulate_one_sided(case, block):
    """Self segments cut into blocks (pairs of different blocks symmetric, diagonal blocks one-sided)
    give the same apply, the same interactions and tables that pass the layout check."""
    from skelfmm_amd.layout import check_layout, split_self_blocks

    _, f, k = case
    L = build_layout(f, max_sources=900)
    one, sym = L.split
    seg = one.e_n[(one.e_flags & SRC_SELF) != 0]
    assert np.any(seg > block), "no self segment to cut"
    assert split_self_blocks(one, sym, L.max_sources, 0) == (one, sym)
    one2, sym2 = split_self_blocks(one, sym, L.max_sources, block)
    assert one2.interactions + 2 * sym2.interactions == L.interactions
    # moved: the pairs of different blocks of every segment longer than the block
    nb = (seg + block - 1) // block
    moved = sum(int(block * (n - block * (i + 1))) for n, b in zip(seg, nb) if n > block for i in range(b - 1))
    assert sym2.interactions == sym.interactions + moved
    assert not np.any(sym2.e_flags & SRC_SELF)
    assert np.all(one2.e_n[(one2.e_flags & SRC_SELF) != 0] <= block)
    L.split = (one2, sym2)
    L.evaluations = one2.interactions + sym2.interactions
    check_layout(L)
    q = np.random.RandomState(4).randn(f.N)
    if f.operator.is_complex:
        q = q + 1j * np.random.RandomState(5).randn(f.N)
    L0 = build_layout(f, max_sources=900)
    assert rel_l2(emulate_apply(L, f.operator.name, k, q, split=True), emulate_apply(L0, f.operator.name, k, q)) < 1e-13


def test_self_blocks_tables_by_hand():
    """A 130-point self segment in blocks of 64: diagonal blocks (64, 64, 2) one-sided with the
    self flag, block pairs (64 x 66 and 64 x 2) symmetric; the box's other entry and a self-flagged
    entry whose charges sit elsewhere stay as they were."""
    from skelfmm_amd.layout import NearLists, _empty_lists, make_items, split_self_blocks

    one = NearLists()
    one.tpi = 64
    i32 = lambda *v: np.array(v, dtype=np.int32)
    one.g_pt, one.g_n, one.g_out, one.g_box, one.g_kind = i32(100, 300), i32(130, 70), i32(100, 1300), i32(3, 5), i32(0, 1)
    one.g_e0, one.g_e1 = np.array([0, 2], np.int64), np.array([2, 3], np.int64)
    # entries: box 3's own points (self), box 7, and box 5's skeleton against its own points with
    # the charges of the other kind (e_q != g_out: not cut)
    one.e_pt, one.e_n, one.e_q = i32(100, 500, 300), i32(130, 40, 70), i32(100, 500, 300)
    one.e_qkind, one.e_box, one.e_flags = i32(0, 0, 0), i32(3, 7, 5), i32(SRC_SELF, 0, SRC_SELF | SRC_NEG)
    one.items, one.item_cost = make_items(one.g_n, one.g_e0, one.g_e1, one.e_n, 1000, 64)
    one.interactions = 130 * 130 + 130 * 40 + 70 * 70
    sym = _empty_lists(512)
    o2, s2 = split_self_blocks(one, sym, 1000, 64)
    grp = lambda P: [tuple(int(getattr(P, k)[g]) for k in ("g_pt", "g_n", "g_out", "g_box", "g_kind", "g_e0", "g_e1"))
                     for g in range(P.g_pt.shape[0])]
    ent = lambda P: [tuple(int(getattr(P, k)[s]) for k in ("e_pt", "e_n", "e_q", "e_box", "e_flags"))
                     for s in range(P.e_pt.shape[0])]
    assert grp(o2) == [(100, 130, 100, 3, 0, 0, 1), (300, 70, 1300, 5, 1, 1, 2), (100, 64, 100, 3, 0, 2, 3),
                       (164, 64, 164, 3, 0, 3, 4), (228, 2, 228, 3, 0, 4, 5)]
    assert ent(o2) == [(500, 40, 500, 7, 0), (300, 70, 300, 5, SRC_SELF | SRC_NEG), (100, 64, 100, 3, SRC_SELF),
                       (164, 64, 164, 3, SRC_SELF), (228, 2, 228, 3, SRC_SELF)]
    assert grp(s2) == [(100, 64, 100, 3, 0, 0, 1), (164, 64, 164, 3, 0, 1, 2)]
    assert ent(s2) == [(164, 66, 164, 3, 0), (228, 2, 228, 3, 0)]
    assert o2.interactions == 130 * 40 + 70 * 70 + 64 * 64 * 2 + 2 * 2 and s2.interactions == 64 * 66 + 64 * 2
    assert o2.interactions + 2 * s2.interactions == one.interactions
    # every item inside its group; blocks at least as long as the segment change nothing
    for P in (o2, s2):
        it = P.items.astype(np.int64)
        assert np.all(it[:, 2] >= P.g_e0[it[:, 0]]) and np.all(it[:, 3] <= P.g_e1[it[:, 0]])
    assert split_self_blocks(one, sym, 1000, 130) == (one, sym)


def test_pair_rows_partition():
    """Every row is one-sided, a pair representative, or the partner of exactly one representative."""
    from skelfmm_amd.layout import pair_rows

    R = np.array([
        [0, 0, 1, 0, 0, 0], [1, 0, 0, 0, 0, 0],    # pair
        [2, 1, 3, 1, 1, 1], [3, 1, 2, 1, 1, 1],    # pair (skeleton SUB)
        [4, 1, 5, 0, 0, 1], [5, 0, 4, 1, 1, 1],    # pair (coarse up / down)
        [6, 0, 6, 0, 0, 0],                        # self: one-sided
        [7, 0, 8, 0, 0, 0],                        # partner missing: one-sided
        [9, 0, 10, 0, 0, 0], [10, 0, 9, 0, 0, 1],  # signs differ: one-sided
        [11, 0, 12, 0, 0, 0], [11, 0, 12, 0, 0, 0], [12, 0, 11, 0, 0, 0],  # duplicate: one-sided
    ], dtype=np.int64)
    one, rep = pair_rows(R, 13)
    assert sorted(map(tuple, rep)) == [(0, 0, 1, 0, 0, 0), (2, 1, 3, 1, 1, 1), (4, 1, 5, 0, 0, 1)]
    assert one.shape[0] + 2 * rep.shape[0] == R.shape[0]
    assert {tuple(r) for r in one} == {tuple(r) for r in R[6:]}


def test_pair_orientation_pads_less():
    """With set sizes, each pair's representative puts the set that leaves fewer idle lanes
    (times sources) in the symmetric kernel's 64-lane rows on the target side: the rows still
    split into one-sided rows plus one representative per pair, and the padded evaluations never
    exceed those of the order-based choice."""
    from skelfmm_amd.layout import SYM_LANES, pair_rows

    rng = np.random.default_rng(3)
    nsets = 400
    size = rng.integers(1, 600, size=nsets)
    A = rng.integers(0, nsets, size=3000)
    B = rng.integers(0, nsets, size=3000)
    keep = A != B
    A, B = A[keep], B[keep]
    pairs = np.unique(np.stack([np.minimum(A, B), np.maximum(A, B)], 1), axis=0)
    rows = np.concatenate([pairs, pairs[:, ::-1]])
    # sets are boxes of kind 0 (box = set index); charge kind 0, sign +
    R = np.stack([rows[:, 0], np.zeros(len(rows), np.int64), rows[:, 1], np.zeros(len(rows), np.int64),
                  np.zeros(len(rows), np.int64), np.zeros(len(rows), np.int64)], 1)
    nt, ns = size[R[:, 0]], size[R[:, 2]]
    one0, rep0 = pair_rows(R, nsets)
    one1, rep1 = pair_rows(R, nsets, sizes=(nt, ns))
    assert one0.shape[0] == one1.shape[0] == 0 and rep0.shape[0] == rep1.shape[0] == pairs.shape[0]
    got = {tuple(sorted((r[0], r[2]))) for r in rep1}
    assert got == {tuple(p) for p in pairs}  # one representative per pair
    pad = lambda n: (n + SYM_LANES - 1) // SYM_LANES * SYM_LANES
    waste = lambda rep: int(np.sum((pad(size[rep[:, 0]]) - size[rep[:, 0]]) * size[rep[:, 2]]))
    assert waste(rep1) <= waste(rep0)
    assert waste(rep1) < waste(rep0)  # sizes this random always leave room


def test_merged_layout_split_emulates_every_instance():
    from skelfmm_amd.layout import merge_layouts

    rng = np.random.default_rng(9)
    builds = [make_build(annulus_points(n, s), ops.LAPLACE_2D, tol=1e-5, leaf_size=40)
              for n, s in ((1500, 1), (3000, 2))]
    L = merge_layouts([build_layout(b, max_sources=600) for b in builds])
    q = rng.normal(size=L.N)
    np.testing.assert_allclose(emulate_apply(L, "laplace_2d", 0.0, q, split=True),
                               emulate_apply(L, "laplace_2d", 0.0, q), rtol=0, atol=1e-11 * np.abs(q).sum())


@pytest.mark.parametrize("case", [c for c in BUILDS if c[0] in ("lap2d", "annulus")], ids=lambda c: c[0])
def test_device_resident_interpolation_matrices(case):
    """T matrices held as slices of result buffers (the GPU build's DeviceProj, here on CPU
    tensors, in shuffled order across two buffers with gaps) gather into the same layout."""
    import copy

    import torch

    from skelfmm_amd.gpu_build import DeviceProj

    _, f, k = case
    L0 = build_layout(f, max_sources=900)
    g = copy.copy(f)
    g.proj = list(f.proj)
    boxes = [b for b in range(f.nboxes) if f.proj[b] is not None and f.proj[b].size > 0]
    assert len(boxes) > 10
    rng = np.random.default_rng(0)
    order = rng.permutation(len(boxes))
    halves = [order[: len(order) // 2], order[len(order) // 2:]]
    for part in halves:
        sizes = [f.proj[boxes[i]].size + 3 for i in part]  # 3 unused entries after each matrix
        buf = torch.zeros(int(sum(sizes)), dtype=torch.float64)
        pos = 0
        for i, sz in zip(part, sizes):
            b = boxes[i]
            p = f.proj[b]
            buf[pos:pos + p.size] = torch.as_tensor(p.ravel())
            g.proj[b] = DeviceProj(buf, pos, p.shape, cplx=False)
            pos += sz
    some = boxes[0]
    np.testing.assert_array_equal(np.asarray(g.proj[some]), f.proj[some])
    L1 = build_layout(g, max_sources=900)
    assert torch.is_tensor(L1.T)
    for b in boxes:
        r, n = f.proj[b].shape
        np.testing.assert_array_equal(L1.T[L1.toff[b]:L1.toff[b] + r * n].numpy().reshape(r, n), f.proj[b])
    L1.T = L1.T.numpy()
    q = np.random.RandomState(3).randn(f.N)
    np.testing.assert_array_equal(emulate_apply(L1, f.operator.name, k, q), emulate_apply(L0, f.operator.name, k, q))


@pytest.mark.parametrize("case", BUILDS, ids=lambda c: c[0])
def test_every_charge_slot_read_is_written_first(case):
    """The apply writes every slot of the charge buffer it reads before reading it, so the
    buffer needs no zero fill per apply: leaf gather, then per upward stage (deepest first)
    the children's skeleton charges into the parents' candidate slots and the boxes'
    skeleton charges; the near field reads candidate and skeleton ranges of its entries."""
    from skelfmm_amd.layout import merge_layouts

    _, f, _ = case
    L = build_layout(f, max_sources=900)
    if case[0] == "annulus":  # merged instances keep the property
        L = merge_layouts([L, build_layout(BUILDS[0][1], max_sources=900)])
    ns = L.nslots
    written = np.zeros(2 * ns, dtype=bool)
    written[L.leaf_slot] = True
    for S in L.up:  # deepest level first, as skf_apply runs them
        assert written[ns + S.gather_src].all(), f"level {S.level}: gather reads unwritten skeleton slots"
        written[S.gather_dst] = True
        for off, bs, r in zip(S.box_off, S.box_bs, S.box_rank):
            assert written[off:off + bs].all(), f"level {S.level}: box candidates read before written"
            written[ns + off:ns + off + r] = True
    lists = [L] + list(L.split or ())
    for P in lists:
        for q0, n in zip(P.e_q, P.e_n):
            assert written[q0:q0 + n].all(), "near field reads an unwritten charge slot"


@pytest.mark.parametrize("case", BUILDS, ids=lambda c: c[0])
def test_xcd_item_order_is_a_padded_permutation(case):
    """The XCD launch order holds every item exactly once plus empty padding items; the items of
    one XCD (workgroups x, x + 8, ...) are runs of the table order of equal cost."""
    _, f, _ = case
    L = build_layout(f, max_sources=700)
    items, costs = order_items(L.items, L.item_cost, "xcd", chunks_per_xcd=2)
    pad = items[:, 3] <= items[:, 2]
    assert np.all(items[pad] == 0) and np.all(costs[pad] == 0)
    real = items[~pad]
    key = lambda a: sorted(map(tuple, a.tolist()))
    assert key(real) == key(L.items)
    assert np.array_equal(np.sort(costs[~pad]), np.sort(L.item_cost))
    table = np.lexsort((L.items[:, 3], L.items[:, 2], L.items[:, 1], L.items[:, 0]))
    rank = {tuple(L.items[i]): r for r, i in enumerate(table)}
    per_xcd = []
    for x in range(N_XCD):
        mine = [rank[tuple(it)] for it in items[x::N_XCD] if it[3] > it[2]]
        # in launch order each XCD walks its runs forward: positions rise within a run
        assert sum(b < a for a, b in zip(mine, mine[1:])) <= 1
        per_xcd.append(sum(int(c) for c, it in zip(costs[x::N_XCD], items[x::N_XCD]) if it[3] > it[2]))
    # runs of equal cost up to one item each: two runs per XCD here
    assert max(per_xcd) - min(per_xcd) <= 2 * 2 * int(L.item_cost.max())
    # padding stops at the last real item: the table is as long as the busiest XCD needs
    lmax = max(int((items[x::N_XCD, 3] > items[x::N_XCD, 2]).sum()) for x in range(N_XCD))
    assert N_XCD * (lmax - 1) < items.shape[0] <= N_XCD * lmax and not pad[-1]
    heavy, hc = order_items(L.items, L.item_cost, "heavy")
    assert np.array_equal(heavy, L.items) and np.array_equal(hc, L.item_cost)
    with pytest.raises(ValueError):
        order_items(L.items, L.item_cost, "spiral")
