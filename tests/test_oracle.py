"""Pins of the CPU oracle: known answers, mathematical identities, cross-checks.

The reference ships no golden vectors (its tests build data through the absent
simpletree package), so the oracle is pinned by (i) the reference tests' own
known answers and identities, restated; (ii) closed forms and independent
series; (iii) finite-difference definitions of the double layers; (iv) the two
apply restatements against each other and against the direct sum; (v) the
committed fixtures in tests/golden.
"""

import math
import os

import numpy as np
import pytest
import torch

from conftest import annulus_points, make_build, rel_l2, sphere_points
from oracle import id as oid
from oracle.batched_apply import BatchedApplyCPU
from oracle.direct import direct_sum_c
from oracle.kernels import OPERATOR_NAMES, direct_sum, kernel_matrix, kernel_matrix_torch
from oracle.serial_apply import SerialApply
from skelfmm_amd import dense
from skelfmm_amd import operators as ops

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- kernels
def test_squared_distance_and_mask_known_answers():
    """test_batched.py:90-107: sq = [[0,4],[1,5]], diagonal mask, FP32 clamp keeps distances positive."""
    xx = torch.tensor([[[0.0, 0.0], [1.0, 0.0]]], dtype=torch.float64)
    yy = torch.tensor([[[0.0, 0.0], [0.0, 2.0]]], dtype=torch.float64)
    K = kernel_matrix_torch("laplace_2d", xx, yy, 0.0, False)
    assert torch.allclose(torch.exp(2 * K), torch.tensor([[[1e-15, 4.0], [1.0, 5.0]]], dtype=torch.float64))
    Ks = kernel_matrix_torch("laplace_3d", xx[..., :2].new_zeros(1, 2, 3) + torch.tensor([0.0, 1.0])[None, :, None],
                             xx.new_zeros(1, 2, 3) + torch.tensor([0.0, 1.0])[None, :, None], 0.0, True)
    assert torch.all(torch.diagonal(Ks[0]) == 0)
    x32 = torch.zeros(1, 2, 2, dtype=torch.float32)
    assert torch.all(torch.isfinite(kernel_matrix_torch("laplace_3d", x32, x32, 0.0, False)))


def test_kernel_diagonals_are_zero():
    """test_util.py:7-19."""
    xx2 = np.array([[0.0, 0.0], [0.25, 0.5], [0.75, 0.1]])
    xx3 = np.array([[0.0, 0.0, 0.0], [0.25, 0.5, 0.1], [0.75, 0.1, 0.2]])
    n3 = xx3 / np.maximum(np.linalg.norm(xx3, axis=1, keepdims=True), 1e-300)
    for name in OPERATOR_NAMES:
        X = xx2 if name.endswith("2d") else xx3
        K = kernel_matrix(name, X, X, 2.0, geom_trg=n3, geom_src=n3)
        assert np.allclose(np.diag(K), 0.0)


def test_closed_form_values():
    x = np.array([[0.0, 0.0, 0.0]])
    y = np.array([[2.0, 0.0, 0.0]])
    assert kernel_matrix("laplace_3d", x, y)[0, 0] == pytest.approx(0.5)
    assert kernel_matrix("laplace_2d", x[:, :2], np.array([[math.e, 0.0]]))[0, 0] == pytest.approx(1.0)
    k = 1.3
    assert kernel_matrix("helmholtz_3d", x, y, k)[0, 0] == pytest.approx(np.exp(2j * k) / 2)


def _series_j0(x, terms=60):
    s, t = 0.0, 1.0
    for m in range(terms):
        if m > 0:
            t *= -(x / 2) ** 2 / (m * m)
        s += t
    return s


def _series_y0(x, terms=60):
    # Y0 = (2/pi)[(ln(x/2) + gamma) J0(x) - sum_{m>=1} (-1)^m H_m (x/2)^{2m} / (m!)^2]
    gamma = 0.57721566490153286
    s, t, H = 0.0, 1.0, 0.0
    for m in range(1, terms):
        t *= -(x / 2) ** 2 / (m * m)
        H += 1.0 / m
        s += H * t
    return (2 / math.pi) * ((math.log(x / 2) + gamma) * _series_j0(x) - s)


def test_helmholtz_2d_against_series():
    for r in (0.3, 1.1, 2.7):
        k = 1.7
        v = kernel_matrix("helmholtz_2d", np.zeros((1, 2)), np.array([[r, 0.0]]), k)[0, 0]
        assert v.real == pytest.approx(_series_j0(k * r), rel=1e-12, abs=1e-13)
        assert v.imag == pytest.approx(_series_y0(k * r), rel=1e-11, abs=1e-12)


def _fd_normal_derivative(G, x, y, n, wrt, h=1e-5):
    if wrt == "source":
        return (G(x, y + h * n) - G(x, y - h * n)) / (2 * h)
    return (G(x + h * n, y) - G(x - h * n, y)) / (2 * h)


@pytest.mark.parametrize("kappa", [0.0, 2.3])
def test_double_layers_are_minus_normal_derivatives(kappa):
    """Code convention (App. C): DL = -d/dn_y G, ADL = -d/dn_x G with G = e^{ikr}/(4 pi r)."""
    rng = np.random.default_rng(4)
    G = lambda a, b: np.exp(1j * kappa * np.linalg.norm(a - b)) / (4 * np.pi * np.linalg.norm(a - b))
    for _ in range(5):
        x, y = rng.normal(size=3), rng.normal(size=3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        dl = "helmholtz_double_layer_3d" if kappa > 0 else "laplace_double_layer_3d"
        adl = "helmholtz_adjoint_double_layer_3d" if kappa > 0 else "laplace_adjoint_double_layer_3d"
        Kd = kernel_matrix(dl, x[None], y[None], kappa, geom_src=n[None])[0, 0]
        Ka = kernel_matrix(adl, x[None], y[None], kappa, geom_trg=n[None])[0, 0]
        fd_s = _fd_normal_derivative(G, x, y, n, "source")
        fd_t = _fd_normal_derivative(G, x, y, n, "target")
        assert abs(Kd - (-fd_s)) < 1e-7 * max(1.0, abs(fd_s))
        assert abs(Ka - (-fd_t)) < 1e-7 * max(1.0, abs(fd_t))


@pytest.mark.parametrize("pair", [("laplace_double_layer_3d", "laplace_adjoint_double_layer_3d", 0.0),
                                  ("helmholtz_double_layer_3d", "helmholtz_adjoint_double_layer_3d", 1.7)])
def test_adjoint_bilinear_identity(pair):
    """test_operators.py:101-137: (K x) . y == x . (K_adj y) at 1e-12."""
    dl, adl, k = pair
    rng = np.random.default_rng(7)
    rs = np.random.RandomState(7)
    src = sphere_points(48, 7)
    trg = rng.normal(size=(32, 3))
    trg /= np.linalg.norm(trg, axis=1, keepdims=True)
    x = rs.randn(48) + (1j * rs.randn(48) if k else 0)
    y = rs.randn(32) + (1j * rs.randn(32) if k else 0)
    K = kernel_matrix(dl, trg, src, k, geom_src=src)
    Ka = kernel_matrix(adl, src, trg, k, geom_trg=src)
    assert np.allclose(np.dot(K @ x, y), np.dot(x, Ka @ y), atol=1e-12, rtol=1e-12)


@pytest.mark.parametrize("op", ops.BUILTIN_OPERATORS, ids=lambda o: o.name)
def test_product_dense_kernels_match_oracle(op):
    """The build's numpy kernels (operators.numpy_func) against the oracle closed forms, incl. r = 0."""
    rng = np.random.default_rng(11)
    d = op.ndim
    X = rng.random((40, d))
    Y = np.vstack([rng.random((30, d)), X[:5]])
    n = rng.normal(size=(75, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    k = 2.1 if op.is_complex else 0.0
    gt = n[:40] if op.geometry_side == "target" else None
    gs = n[:35] if op.geometry_side == "source" else None
    A = op.numpy_matrix(X, Y, geom_trg=gt, geom_src=gs, param=k)
    B = kernel_matrix(op.name, X, Y, k, geom_trg=gt, geom_src=gs)
    assert np.allclose(A, B, rtol=1e-13, atol=1e-13)
    far = op.numpy_matrix(X, Y[:30] + 3.0, geom_trg=gt, geom_src=None if gs is None else gs[:30], param=k,
                          assume_not_self_interaction=True)
    dflt = op.numpy_matrix(X, Y[:30] + 3.0, geom_trg=gt, geom_src=None if gs is None else gs[:30], param=k)
    assert np.allclose(far, dflt)


def test_c_direct_sum_matches_numpy():
    rng = np.random.default_rng(2)
    for name in OPERATOR_NAMES:
        d = 2 if name.endswith("2d") else 3
        X, Y = rng.random((50, d)), rng.random((70, d))
        Y[:3] = X[:3]
        n = sphere_points(70, 3)
        q = rng.normal(size=(70, 2))
        if name in ("helmholtz_2d", "helmholtz_3d", "helmholtz_double_layer_3d", "helmholtz_adjoint_double_layer_3d"):
            q = q + 1j * rng.normal(size=(70, 2))
        gt = n[:50] if "adjoint" in name else None
        gs = n if ("double" in name and "adjoint" not in name) else None
        a = direct_sum(name, X, Y, q, 1.9, geom_trg=gt, geom_src=gs)
        b = direct_sum_c(name, X, Y, q, 1.9, geom_trg=gt, geom_src=gs)
        assert rel_l2(b, a) < 1e-13, name


# --------------------------------------------------------------------- ID
def test_pivoted_qr_id_reconstructs():
    """test_util.py:115-129."""
    A = np.random.RandomState(7).randn(18, 8)
    rank, idx, proj = oid.skeleton_id(A.copy(), 1e-10)
    assert 0 < rank <= 8 and idx.shape == (8,) and proj.shape == (rank, 8 - rank)
    if rank < 8:
        approx = A[:, idx[:rank]] @ proj
        assert np.linalg.norm(approx - A[:, idx[rank:]]) / np.linalg.norm(A[:, idx[rank:]]) < 1e-8


def test_id_eps_and_fixed_rank_agree():
    """test_util.py:60-92 on a proxy matrix."""
    rng = np.random.default_rng(0)
    xb = rng.random((8, 2)) * 0.4 - 0.2
    K = oid.proxy_matrix(ops.LAPLACE_2D, xb, np.zeros(2), 0.5, 24)
    r1, i1, p1 = oid.skeleton_id(K.copy(), 1e-6)
    r2, i2, p2 = oid.skeleton_id(K.copy(), r1)
    assert r1 == r2 and np.array_equal(i1, i2) and p1.shape == p2.shape


def test_low_rank_proxy_id_reconstructs_far_field():
    """Skeleton charges reproduce a box's field outside its proxy surface to ~tol."""
    rng = np.random.default_rng(5)
    xb = rng.random((200, 3)) * 0.25
    c = np.full(3, 0.125)
    K = oid.proxy_matrix(ops.LAPLACE_3D, xb, c, 0.25, 512)
    r, idx, T = oid.skeleton_id(K.copy(), 1e-8)
    assert r < 200
    q = rng.normal(size=200)
    qs = q[idx[:r]] + T @ q[idx[r:]]
    far = c + 2.0 * rng.normal(size=(50, 3))
    a = kernel_matrix("laplace_3d", far, xb) @ q
    b = kernel_matrix("laplace_3d", far, xb[idx[:r]]) @ qs
    assert rel_l2(b, a) < 1e-6


def test_build_id_equals_stacked_reference_form():
    """The build factors K_BP^T once; the reference stacks [K_BP^T; K_BP^T] (util.py:90-101)."""
    from skelfmm_amd.build import build_proxy_matrix, skeleton_id_pivoted_qr

    rng = np.random.default_rng(3)
    xb = rng.random((120, 3)) * 0.2
    c = np.full(3, 0.1)
    K1 = build_proxy_matrix(ops.LAPLACE_3D, xb, c, 0.2, 400)
    r1, i1, p1 = skeleton_id_pivoted_qr(K1.copy(), 1e-7)
    K2 = oid.proxy_matrix(ops.LAPLACE_3D, xb, c, 0.2, 400)
    r2, i2, p2 = oid.skeleton_id(K2.copy(), 1e-7)
    assert r1 == r2 and np.array_equal(i1, i2)
    assert np.allclose(p1, p2, rtol=1e-8, atol=1e-10)


# ------------------------------------------------------------ apply oracles
def _problems():
    rng = np.random.default_rng(0)
    yield "lap2d", rng.random((2500, 2)), ops.LAPLACE_2D, 0.0, None, 50
    yield "lap2d_annulus", annulus_points(3000, 1), ops.LAPLACE_2D, 0.0, None, 60
    yield "lap3d", rng.random((8000, 3)), ops.LAPLACE_3D, 0.0, None, 120
    yield "helm3d", rng.random((2500, 3)), ops.HELMHOLTZ_3D, 4.0, None, 80
    v = sphere_points(3000, 2)
    yield "ldl", v, ops.LAPLACE_DOUBLE_LAYER_3D, 0.0, v, 80
    yield "hadl", v, ops.HELMHOLTZ_ADJOINT_DOUBLE_LAYER_3D, 3.0, v, 80


@pytest.mark.parametrize("case", list(_problems()), ids=lambda c: c[0])
def test_serial_and_batched_oracles_agree_and_approximate_direct_sum(case):
    """Batched == serial (test_batched.py:403-466, tightened to 1e-12); FMM vs dense below tol (test_fmm.py)."""
    _, X, op, k, g, leaf = case
    tol = 1e-6
    fmm = make_build(X, op, kappa=k, tol=tol, leaf_size=leaf, geometry=g)
    q = np.random.RandomState(1).randn(X.shape[0])
    ys = SerialApply(fmm).matvec(q)
    yb = BatchedApplyCPU(fmm, bytes_budget=1e8).matvec(q).numpy()
    assert rel_l2(yb, ys) < 1e-12
    gt = g if op.geometry_side == "target" else None
    gs = g if op.geometry_side == "source" else None
    yd = direct_sum_c(op.name, X, X, q, k, geom_trg=gt, geom_src=gs)
    assert rel_l2(ys, yd) < 10 * tol
    idx = np.random.default_rng(0).choice(X.shape[0], 64, replace=False)
    assert np.linalg.norm(ys[idx] - yd[idx], np.inf) / np.linalg.norm(yd[idx], np.inf) < 10 * tol


def test_oracle_detects_planted_mistakes():
    """Planted mistakes in the oracle (dropped list, transposed T, swapped child order) break the direct-sum check."""
    X = np.random.default_rng(1).random((3000, 2))
    X[:600] *= 0.1  # clustered corner -> coarse lists are non-empty
    fmm = make_build(X, ops.LAPLACE_2D, tol=1e-8, leaf_size=40)
    assert fmm.u_from_qskel_list.shape[0] > 0
    q = np.random.RandomState(2).randn(X.shape[0])
    yd = direct_sum_c("laplace_2d", X, X, q)
    assert rel_l2(SerialApply(fmm).matvec(q), yd) < 1e-6

    saved = fmm.uskel_from_qskel_list
    fmm.uskel_from_qskel_list = fmm.tree_uskel_from_qskel_list  # drop the leaf colleague pairs
    assert rel_l2(SerialApply(fmm).matvec(q), yd) > 1e-3
    fmm.uskel_from_qskel_list = saved

    box = next(b for b in range(fmm.nboxes) if fmm.proj[b] is not None and fmm.proj[b].shape[0] > 1
               and fmm.proj[b].shape[1] > 1)
    saved = fmm.proj[box]
    fmm.proj[box] = -saved  # wrong sign of one interpolation matrix
    assert rel_l2(SerialApply(fmm).matvec(q), yd) > 1e-6
    fmm.proj[box] = saved

    P = next(b for b in range(fmm.nboxes) if fmm.child_src[b] is not None and fmm.child_src[b].shape[0] > 2)
    saved = fmm.child_src[P]
    fmm.child_src[P] = saved[::-1].copy()  # wrong parent/child slot map
    assert rel_l2(SerialApply(fmm).matvec(q), yd) > 1e-6
    fmm.child_src[P] = saved


def test_single_precision_batched_oracle_is_close():
    X = np.random.default_rng(1).random((2000, 3))
    fmm = make_build(X, ops.HELMHOLTZ_3D, kappa=2.0, tol=1e-5, leaf_size=60)
    q = np.random.default_rng(0).random(X.shape[0])
    y64 = BatchedApplyCPU(fmm, bytes_budget=1e8).matvec(q).numpy()
    y32 = BatchedApplyCPU(fmm, use_single_precision=True, bytes_budget=1e8).matvec(q.astype(np.float32)).numpy()
    assert y32.dtype == np.complex64
    assert rel_l2(y32.astype(np.complex128), y64) < 5e-4


# -------------------------------------------------------------- fixtures
@pytest.mark.parametrize("fname", sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz")) if os.path.isdir(GOLDEN) else [])
def test_golden_fixture(fname):
    """Committed fixtures: kernels exactly, and the apply of a rebuilt tree within the stated tolerance."""
    z = np.load(os.path.join(GOLDEN, fname), allow_pickle=False)
    name = str(z["operator"])
    kappa = float(z["kappa"])
    X, q = z["points"], z["charges"]
    geom = z["normals"] if "normals" in z.files else None
    gt = geom if name in ("laplace_adjoint_double_layer_3d", "helmholtz_adjoint_double_layer_3d") else None
    gs = geom if geom is not None and gt is None else None
    yd = direct_sum_c(name, X, X, q, kappa, geom_trg=gt, geom_src=gs)
    assert rel_l2(yd, z["direct"]) < 1e-13
    op = {o.name: o for o in ops.BUILTIN_OPERATORS}[name]
    fmm = make_build(X, op, kappa=kappa, tol=float(z["tol"]), leaf_size=int(z["leaf_size"]), geometry=geom)
    y = BatchedApplyCPU(fmm, bytes_budget=1e8).matvec(q).numpy()
    assert rel_l2(y, z["apply"]) < 10 * float(z["tol"])
    assert rel_l2(y, yd) < 10 * float(z["tol"])
