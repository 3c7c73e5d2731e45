"""Serialized built operator (CPU): round trip of every layout array, version and checksum checks.

A layout read back must be the same layout: every array equal in value and dtype, every
scalar equal, and the emulated apply (the numpy statement of the device apply) identical
on the round-tripped layout. A file of another format, a newer version or with a changed
byte must be rejected.
"""

import io
import json

import numpy as np
import pytest

from conftest import make_build, rel_l2, sphere_points
from layout_emulator import emulate_apply
from skelfmm_amd import operators as ops
from skelfmm_amd.layout import build_layout, resplit
from skelfmm_amd.serialize import FORMAT, VERSION, load_layout, save_layout


def _cases():
    rng = np.random.default_rng(7)
    yield "lap3d", make_build(rng.random((6000, 3)), ops.LAPLACE_3D, tol=1e-6, leaf_size=80), 0.0, None
    v = sphere_points(2500, 2)
    yield "hdl", make_build(v, ops.HELMHOLTZ_DOUBLE_LAYER_3D, tol=1e-5, leaf_size=60, geometry=v, kappa=3.0), 3.0, v
    yield "lap2d", make_build(rng.random((3000, 2)), ops.LAPLACE_2D, tol=1e-6, leaf_size=40), 0.0, None


CASES = list(_cases())


def _roundtrip(f, kappa, geom, **kw):
    L = build_layout(f, max_sources=2048)
    buf = io.BytesIO()
    save_layout(buf, L, operator_name=f.operator.name, kappa=kappa, tol=f.tol, leaf_size=f.tree.leaf_size,
                points=f.tree.XX, geometry=geom, extra={"note": "test"},
                extra_arrays={"rank_list": f.rank_list, "bs_list": f.bs_list}, **kw)
    buf.seek(0)
    return L, buf


def _same(a, b, where):
    assert (a is None) == (b is None), where
    if a is None:
        return
    assert a.dtype == b.dtype and a.shape == b.shape, where
    assert np.array_equal(a, b), where


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_roundtrip_restores_every_array(case):
    _, f, kappa, geom = case
    L, buf = _roundtrip(f, kappa, geom)
    rec = load_layout(buf)
    M = rec.layout
    assert rec.header["format"] == FORMAT and rec.header["version"] == VERSION
    assert rec.header["operator"] == f.operator.name and rec.header["kappa"] == kappa
    assert rec.header["extra"] == {"note": "test"}
    for k in ("ndim", "N", "nboxes", "root_level", "nslots", "interactions", "evaluations", "max_sources"):
        assert getattr(M, k) == getattr(L, k), k
    for k in ("off", "slot_point", "slot_box", "points", "normals", "centers", "leaf_slot", "leaf_user", "T", "toff",
              "items", "item_cost", "g_pt", "g_n", "g_out", "g_box", "g_kind", "g_e0", "g_e1", "e_pt", "e_n", "e_q",
              "e_qkind", "e_box", "e_flags"):
        _same(getattr(L, k), getattr(M, k), k)
    for name in ("up", "down"):
        assert len(getattr(L, name)) == len(getattr(M, name))
        for S, T in zip(getattr(L, name), getattr(M, name)):
            assert S.level == T.level
            for k in ("boxes", "box_off", "box_bs", "box_rank", "box_toff", "gather_dst", "gather_src", "items"):
                _same(getattr(S, k), getattr(T, k), f"{name}{S.level}.{k}")
    assert (L.split is None) == (M.split is None)
    for P, Q in zip(L.split or (), M.split or ()):
        assert P.tpi == Q.tpi and P.interactions == Q.interactions
        for k in P.FIELDS + ("items", "item_cost"):
            _same(getattr(P, k), getattr(Q, k), "split." + k)
    np.testing.assert_array_equal(rec.points, f.tree.XX)
    _same(None if geom is None else np.ascontiguousarray(geom), rec.geometry, "geometry")
    np.testing.assert_array_equal(rec.extra["rank_list"], f.rank_list)
    np.testing.assert_array_equal(rec.extra["bs_list"], f.bs_list)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_roundtrip_apply_is_identical(case):
    _, f, kappa, geom = case
    L, buf = _roundtrip(f, kappa, geom)
    M = load_layout(buf).layout
    q = np.random.RandomState(1).randn(f.N)
    if f.operator.is_complex:
        q = q + 1j * np.random.RandomState(2).randn(f.N)
    y0 = emulate_apply(L, f.operator.name, kappa, q, split=L.split is not None)
    y1 = emulate_apply(M, f.operator.name, kappa, q, split=M.split is not None)
    assert np.array_equal(y0, y1)
    # the loaded layout re-splits like a fresh one
    resplit(M, 512)
    y2 = emulate_apply(M, f.operator.name, kappa, q)
    assert rel_l2(y2, y0) < 1e-13


def test_roundtrip_of_self_blocks():
    """A layout whose self segments were cut into symmetric blocks saves and loads like any other:
    the loaded tables pass the layout check, apply identically, and cut again (block 0 and a
    block at least as long as the diagonal blocks) change nothing."""
    from skelfmm_amd.layout import split_self_blocks

    _, f, kappa, geom = CASES[0]
    L = build_layout(f, max_sources=2048)
    L.split = split_self_blocks(*L.split, L.max_sources, 32)
    L.evaluations = L.split[0].interactions + L.split[1].interactions
    buf = io.BytesIO()
    save_layout(buf, L, operator_name=f.operator.name, kappa=kappa, tol=f.tol, leaf_size=f.tree.leaf_size,
                points=f.tree.XX)
    buf.seek(0)
    M = load_layout(buf).layout
    for P, Q in zip(L.split, M.split):
        assert P.tpi == Q.tpi and P.interactions == Q.interactions
        for k in P.FIELDS + ("items", "item_cost"):
            _same(getattr(P, k), getattr(Q, k), "split." + k)
    q = np.random.RandomState(1).randn(f.N)
    y0 = emulate_apply(L, f.operator.name, kappa, q, split=True)
    assert np.array_equal(emulate_apply(M, f.operator.name, kappa, q, split=True), y0)
    assert split_self_blocks(*M.split, M.max_sources, 0) == M.split
    assert split_self_blocks(*M.split, M.max_sources, 32) == M.split
    assert rel_l2(y0, emulate_apply(M, f.operator.name, kappa, q)) < 1e-13


def _rewrite(buf, mutate_header=None, mutate_arrays=None):
    buf.seek(0)
    with np.load(buf, allow_pickle=False) as z:
        arrays = {k: z[k].copy() for k in z.files}
    header = json.loads(bytes(arrays["__header__"]).decode())
    if mutate_header:
        mutate_header(header)
    if mutate_arrays:
        mutate_arrays(arrays)
    arrays["__header__"] = np.frombuffer(json.dumps(header).encode(), dtype=np.uint8)
    out = io.BytesIO()
    np.savez(out, **arrays)
    out.seek(0)
    return out


def test_rejects_other_format_and_newer_version():
    _, f, kappa, geom = CASES[0]
    _, buf = _roundtrip(f, kappa, geom)
    with pytest.raises(ValueError, match="format"):
        load_layout(_rewrite(buf, mutate_header=lambda h: h.update(format="something.else")))
    with pytest.raises(ValueError, match="version"):
        load_layout(_rewrite(buf, mutate_header=lambda h: h.update(version=VERSION + 1)))
    with pytest.raises(ValueError, match="no header"):
        out = io.BytesIO()
        np.savez(out, points=np.zeros(3))
        out.seek(0)
        load_layout(out)


def test_detects_a_changed_byte():
    _, f, kappa, geom = CASES[0]
    _, buf = _roundtrip(f, kappa, geom)

    def flip(arrays):
        a = arrays["layout/T"].view(np.uint8)
        a[a.shape[0] // 2] ^= 1

    bad = _rewrite(buf, mutate_arrays=flip)
    with pytest.raises(ValueError, match="checksum mismatch in 'layout/T'"):
        load_layout(bad)
    bad.seek(0)
    load_layout(bad, verify=False)  # still readable when the check is off


def _plant(key, mutate):
    """A copy of the first case's file with ``mutate(array)`` applied to ``key`` and the CRC
    recomputed, as a hand-made or inconsistent file would carry it."""
    from skelfmm_amd.serialize import _crc

    _, f, kappa, geom = CASES[0]
    _, buf = _roundtrip(f, kappa, geom)

    planted = {}

    def arrays_fn(arrays):
        arrays[key] = arrays[key].copy()
        mutate(arrays[key])
        planted["crc"] = _crc(arrays[key])

    def header_fn(header):
        header["crc32"][key] = planted["crc"]

    # _rewrite applies the array change first, then the header change
    buf.seek(0)
    with np.load(buf, allow_pickle=False) as z:
        arrays_fn({key: z[key]})
    return _rewrite(buf, mutate_header=header_fn, mutate_arrays=arrays_fn)


def _set_last(v):
    def f(a):
        a.reshape(-1)[-1] = v
    return f


@pytest.mark.parametrize("key,mutate", [
    ("layout/leaf_slot", _set_last(10 ** 8)),
    ("layout/g_pt", _set_last(-1)),
    ("layout/e_q", _set_last(2 ** 30)),
    ("layout/e_box", _set_last(10 ** 6)),
    ("layout/items", _set_last(10 ** 6)),            # entry range past the group's entries
    ("layout/g_e1", _set_last(10 ** 7)),
    ("up/1/gather_src", _set_last(10 ** 8)),
    ("up/0/box_toff", _set_last(10 ** 12)),         # T block past the end of T
    ("up/0/items", _set_last(-5)),
    ("split/1/e_n", _set_last(10 ** 7)),
    ("layout/toff", _set_last(-3)),
], ids=lambda v: v if isinstance(v, str) else "")
def test_rejects_out_of_range_indices(key, mutate):
    """The CRC sits in the same file, so a consistent-looking file with an index that leaves
    its array must still be refused before any table reaches the device."""
    with pytest.raises(ValueError, match="inconsistent box layout"):
        load_layout(_plant(key, mutate))


def test_check_layout_accepts_every_built_layout():
    from skelfmm_amd.layout import check_layout

    for _, f, _, _ in CASES:
        L = build_layout(f, max_sources=512)
        assert check_layout(L) is L


@pytest.mark.gpu
@pytest.mark.parametrize("single", [False, True])
def test_saved_operator_applies_like_the_original(tmp_path, single):
    """SkelFMM.save -> SkelFMM.load: same apply on the GPU, in either precision form."""
    import torch

    from skelfmm_amd import SkelFMM

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    X = np.random.default_rng(5).random((30000, 3))
    fmm = SkelFMM(X, ops.LAPLACE_3D, tol=1e-6, leaf_size=200, use_single_precision=single)
    x = np.random.RandomState(1).randn(X.shape[0])
    y0 = fmm.apply(x)
    path = tmp_path / "laplace3d.npz"
    fmm.save(path)
    g = SkelFMM.load(path, use_single_precision=single)
    assert g.kernel_backend == "hip" and g.nlevels == fmm.nlevels and g.rank_max == fmm.rank_max
    y1 = g.apply(x)
    print(f"single={single}: reload rel-l2 {rel_l2(y1, y0):.2e}", flush=True)
    assert rel_l2(y1, y0) < (1e-6 if single else 1e-14)
    other = SkelFMM.load(path, use_single_precision=not single).apply(x)
    assert rel_l2(other, y0) < 3e-6
    assert g.relerr_fmm_apply_check(x) < 1e-5


@pytest.mark.gpu
def test_saved_double_layer_keeps_its_geometry(tmp_path):
    import torch

    from skelfmm_amd import SkelFMM

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    v = sphere_points(8000, 3)
    fmm = SkelFMM(v, ops.HELMHOLTZ_DOUBLE_LAYER_3D, geometry=v, param=4.0, tol=1e-6, leaf_size=100)
    x = np.random.RandomState(1).randn(v.shape[0]) + 1j * np.random.RandomState(2).randn(v.shape[0])
    y0 = fmm.apply(x)
    fmm.save(tmp_path / "hdl.npz")
    g = SkelFMM.load(tmp_path / "hdl.npz")
    np.testing.assert_array_equal(g.geometry, v)
    assert g.param == 4.0
    assert rel_l2(g.apply(x), y0) < 1e-14
