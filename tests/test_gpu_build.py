"""GPU tests of the batched build kernels and the learned rank predictor.

* ``skf_proxy_matrices`` against the host ``build_proxy_matrix`` (FP64, 1e-13);
* ``skf_pqr_batched`` + ``skf_trsm_batched`` against scipy's pivoted QR on
  matrices with a clear singular-value gap (same rank, ID error below tol);
* GPU adaptive build: apply error against the FP64 direct sum, ranks close to
  the host build's;
* sketched build with the predicted rank forced to 1: every box falls back and
  the skeletons equal the adaptive GPU build's exactly;
* MLP forward / backward / Adam kernels against ``oracle.mlp`` (FP32 bars
  stated per test), training on the GPU, and the trained network's gain over a
  rank predictor without it (per-(operator, tolerance, level) mean rank);
* learned build accuracy and its fallback rate.
"""

import ctypes

import numpy as np
import pytest
import torch
from scipy.linalg import qr

from conftest import annulus_points, rel_l2, sphere_points
from oracle import mlp
from oracle.direct import direct_sum_c
from oracle.kernels import GEOM_SIDE
from skelfmm_amd import hip
from skelfmm_amd import learned_rank as lr
from skelfmm_amd import operators as ops
from skelfmm_amd.build import RecursiveSkeletonization, build_proxy_matrix
from skelfmm_amd.tree import BalancedTree

pytestmark = pytest.mark.gpu


def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    return torch.device("cuda", 0)


def _build(x, op, kappa, tol, leaf, device, rank_model=None, geometry=None):
    tree = BalancedTree(x, leaf_size=leaf)
    f = RecursiveSkeletonization(tree, op, kappa=kappa, tol=tol, geometry=geometry)
    f.skel_tree(device=device, rank_model=rank_model)
    f.setup_lists()
    return f


def _apply_err(fmm, x, seed=1, geometry=None):
    from skelfmm_amd.batched import SkelFMMBatched

    dev = SkelFMMBatched(fmm.tree, fmm.operator, kappa=fmm.kappa).build_from_copy(fmm)
    rng = np.random.RandomState(seed)
    q = rng.randn(x.shape[0])
    if fmm.operator.is_complex:
        q = q + 1j * rng.randn(x.shape[0])
    y = dev.matvec(torch.as_tensor(q).to(_gpu())).cpu().numpy()
    idx = np.random.default_rng(0).choice(x.shape[0], 600, replace=False)
    side = GEOM_SIDE.get(fmm.operator.name)
    ref = direct_sum_c(fmm.operator.name, x[idx], x, q, kappa=fmm.kappa,
                       geom_trg=geometry[idx] if side == "target" else None,
                       geom_src=geometry if side == "source" else None)
    return rel_l2(y[idx], ref)


# ----------------------------------------------------------- proxy matrices
@pytest.mark.parametrize("opname,kappa", [("laplace_3d", 0.0), ("helmholtz_3d", 7.0),
                                          ("helmholtz_double_layer_3d", 5.0), ("laplace_2d", 0.0)])
def test_proxy_matrices_match_host(opname, kappa):
    from skelfmm_amd.gpu_build import GPUSkeletonizer

    _gpu()
    op = _op(opname)
    rng = np.random.default_rng(0)
    d = 2 if op.ndim == 2 else 3
    X = sphere_points(3000) if op.requires_geometry else rng.random((3000, d))
    geom = X if op.requires_geometry else None
    g = GPUSkeletonizer(op, kappa, 1e-6, X, geom)
    cands = [rng.choice(3000, k, replace=False) for k in (50, 170, 9)]
    centers = np.array([X[c].mean(axis=0) for c in cands])
    A, _, a_off, m, ncols, _ = g._proxy(cands, centers, 0.25, 300)
    torch.cuda.synchronize()
    Ah = A.cpu().numpy()
    if op.is_complex:
        Ah = Ah.view(np.complex128)
    for b, c in enumerate(cands):
        ref = build_proxy_matrix(op, X[c], centers[b], 0.25, 300, geom=None if geom is None else geom[c], param=kappa)
        got = Ah[a_off[b]:a_off[b] + m * len(c)].reshape(m, len(c))
        assert got.shape == ref.shape
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


# ------------------------------------------------------------------- QR
def _gap_matrix(m, n, k, rng, cplx=False):
    p = min(m, n)
    g = lambda r, c: rng.normal(size=(r, c)) + (1j * rng.normal(size=(r, c)) if cplx else 0)
    U = np.linalg.qr(g(m, p))[0]
    V = np.linalg.qr(g(n, p))[0]
    s = np.concatenate([np.logspace(0, -3, k), np.full(p - k, 1e-11)])
    return (U * s) @ V.conj().T


@pytest.mark.parametrize("cplx", [False, True])
def test_batched_pivoted_qr_matches_scipy(cplx):
    dev = _gpu()
    L = hip.lib()
    rng = np.random.default_rng(0)
    shapes = [(200, 60, 17), (150, 150, 40), (90, 250, 33), (64, 8, 8)]
    mats = [_gap_matrix(m, n, k, rng, cplx) for m, n, k in shapes]
    es = 2 if cplx else 1
    a_off = np.concatenate(([0], np.cumsum([A.size for A in mats])[:-1])).astype(np.int64)
    flat = np.concatenate([A.ravel() for A in mats])
    buf = torch.as_tensor(flat.view(np.float64) if cplx else flat, device=dev).contiguous()
    nb = len(mats)
    # every device array is held in a variable until the kernels have run: a temporary freed
    # before the launch would hand its memory to the next one
    t = lambda a, dt=torch.int32: torch.as_tensor(np.asarray(a), dtype=dt, device=dev)
    ms, ns = [s[0] for s in shapes], [s[1] for s in shapes]
    d_aoff, d_m, d_n, d_k = t(a_off, torch.int64), t(ms), t(ns), t(np.minimum(ms, ns))
    rank, conv = t(np.zeros(nb)), t(np.zeros(nb))
    p_off = np.concatenate(([0], np.cumsum(ns)[:-1])).astype(np.int64)
    d_poff = t(p_off, torch.int64)
    perm = t(np.zeros(sum(ns)))
    rdiag = torch.zeros(sum(ns), dtype=torch.float64, device=dev)
    tol = 1e-7
    st = hip.stream_handle(dev)
    hip.check_build(L.skf_pqr_batched(hip.ptr(buf), hip.ptr(d_aoff), None, hip.ptr(d_m), hip.ptr(d_n), hip.ptr(d_k), tol,
                                      hip.ptr(rank), hip.ptr(conv), hip.ptr(perm), hip.ptr(d_poff), hip.ptr(rdiag), nb,
                                      max(ms), max(ns), int(cplx), st))
    r_h = rank.cpu().numpy()
    assert np.all(conv.cpu().numpy() == 1)
    t_sizes = [int(r) * (n - int(r)) for r, n in zip(r_h, ns)]
    t_off = np.concatenate(([0], np.cumsum(t_sizes)[:-1])).astype(np.int64)
    d_toff = t(t_off, torch.int64)
    T = torch.zeros(max(1, sum(t_sizes)) * es, dtype=torch.float64, device=dev)
    hip.check_build(L.skf_trsm_batched(hip.ptr(buf), hip.ptr(d_aoff), None, hip.ptr(d_n), hip.ptr(rank), hip.ptr(T),
                                       hip.ptr(d_toff), nb, max(ns), int(cplx), st))
    torch.cuda.synchronize()
    Th = T.cpu().numpy()
    Th = Th.view(np.complex128) if cplx else Th
    perm_h = perm.cpu().numpy()
    for b, (A, (m, n, k)) in enumerate(zip(mats, shapes)):
        R = qr(A, mode="r", pivoting=True)[0]
        d = np.abs(np.diag(R))
        assert r_h[b] == np.sum(d > tol * d[0]) == k
        piv = perm_h[p_off[b]:p_off[b] + n]
        assert sorted(piv.tolist()) == list(range(n))
        r = int(r_h[b])
        Tb = Th[t_off[b]:t_off[b] + r * (n - r)].reshape(r, n - r)
        if r < n:
            E = A[:, piv[r:]] - A[:, piv[:r]] @ Tb
            assert np.linalg.norm(E, 2) <= 1e-8 * np.linalg.norm(A, 2)


@pytest.mark.parametrize("cplx", [False, True])
def test_id_residual_kernel_matches_host(cplx):
    """skf_id_residual against the host acceptance rule's residual on random sketches."""
    from skelfmm_amd.gpu_build import accept_sketch

    dev = _gpu()
    L = hip.lib()
    rng = np.random.default_rng(7)
    c, lda = 16, 90
    ns = np.array([90, 40, 7, 1, 60], dtype=np.int64)
    rk = np.array([30, 40, 0, 1, 12], dtype=np.int64)
    nb = len(ns)
    g = lambda *sh: rng.normal(size=sh) + (1j * rng.normal(size=sh) if cplx else 0)
    Z = g(nb, c, lda)
    perms = [rng.permutation(int(n)) for n in ns]
    Ts = [g(int(r), int(n - r)) for r, n in zip(rk, ns)]
    p_off = np.concatenate(([0], np.cumsum(ns)[:-1])).astype(np.int64)
    sizes = rk * (ns - rk)
    t_off = np.concatenate(([0], np.cumsum(sizes)[:-1])).astype(np.int64)
    Tflat = np.concatenate([T.ravel() for T in Ts] + [np.zeros(1, dtype=Z.dtype)])
    as_dev = lambda a: torch.as_tensor(np.ascontiguousarray(a).view(np.float64) if cplx else a, device=dev)
    d_Z, d_T = as_dev(Z.ravel()), as_dev(Tflat)
    i32 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int32), device=dev)
    i64 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int64), device=dev)
    d_zoff, d_n, d_r, d_perm, d_poff, d_toff = (i64(np.arange(nb) * c * lda), i32(ns), i32(rk),
                                                 i32(np.concatenate(perms)), i64(p_off), i64(t_off))
    res = torch.empty(nb, dtype=torch.float64, device=dev)
    zn = torch.empty(nb, dtype=torch.float64, device=dev)
    d_lda = i32(np.full(nb, lda))
    hip.check_build(L.skf_id_residual(hip.ptr(d_Z), hip.ptr(d_zoff), hip.ptr(d_lda), c, hip.ptr(d_n), hip.ptr(d_r),
                                      hip.ptr(d_perm), hip.ptr(d_poff), hip.ptr(d_T), hip.ptr(d_toff), nb, int(cplx),
                                      hip.ptr(res), hip.ptr(zn), hip.stream_handle(dev)))
    torch.cuda.synchronize()
    res_h, zn_h = res.cpu().numpy(), zn.cpu().numpy()
    for b in range(nb):
        n, r = int(ns[b]), int(rk[b])
        Zb = Z[b, :, :n]
        want_zn = np.max(np.linalg.norm(Zb, axis=0))
        E = Zb[:, perms[b][r:]] - (Zb[:, perms[b][:r]] @ Ts[b] if r else 0)
        want_res = np.max(np.linalg.norm(E, axis=0)) if n > r else 0.0
        assert zn_h[b] == pytest.approx(want_zn, rel=1e-13)
        assert res_h[b] == pytest.approx(want_res, rel=1e-12, abs=1e-300)
        # the GPU numbers decide exactly as the host rule does
        thr = res_h[b] / (1e-6 * want_zn) * 1.0001
        assert accept_sketch(Zb, (r, perms[b], Ts[b]), n, 1e-6, cap=n, accept=thr)
        assert n == r or not accept_sketch(Zb, (r, perms[b], Ts[b]), n, 1e-6, cap=n, accept=thr / 1.001)


# --------------------------------------------------------------- builds
BUILD_CASES = [
    ("laplace_3d", "cube", 0.0, 1e-6, 64),
    ("helmholtz_3d", "cube", 12.0, 1e-6, 64),
    ("laplace_2d", "annulus", 0.0, 1e-7, 64),
    ("helmholtz_double_layer_3d", "sphere", 6.0, 1e-5, 64),
]


def _cloud(kind, n):
    if kind == "cube":
        return np.random.default_rng(0).random((n, 3)), None
    if kind == "annulus":
        return annulus_points(n), None
    x = sphere_points(n)
    return x, x


def _op(name):
    return next(o for o in ops.BUILTIN_OPERATORS if o.name == name)


@pytest.mark.parametrize("case", BUILD_CASES, ids=[c[0] for c in BUILD_CASES])
def test_gpu_adaptive_build_accuracy(case):
    _gpu()
    name, kind, kappa, tol, leaf = case
    op = _op(name)
    x, geom = _cloud(kind, 20000)
    f_gpu = _build(x, op, kappa, tol, leaf, "gpu", geometry=geom)
    f_cpu = _build(x, op, kappa, tol, leaf, "cpu", geometry=geom)
    assert f_gpu.gpu_stats["cpu"] == 0 and f_gpu.gpu_stats["adaptive"] > 0
    boxes = np.nonzero(f_cpu.rank_list)[0]
    dr = np.abs(f_gpu.rank_list[boxes] - f_cpu.rank_list[boxes])
    print(f"{name}: mean |rank diff| {dr.mean():.3f}, max {dr.max()}, "
          f"sum ranks gpu {f_gpu.rank_list.sum()} cpu {f_cpu.rank_list.sum()}")
    # ties at the cutoff |R_ii| ~ tol |R_00| round differently; parents inherit the difference
    assert dr.max() <= max(3, 0.05 * f_cpu.rank_list.max())
    assert abs(int(f_gpu.rank_list.sum()) - int(f_cpu.rank_list.sum())) <= 0.02 * f_cpu.rank_list.sum()
    e_gpu = _apply_err(f_gpu, x, geometry=geom)
    e_cpu = _apply_err(f_cpu, x, geometry=geom)
    print(f"{name}: apply error gpu-build {e_gpu:.2e} cpu-build {e_cpu:.2e} (tol {tol:.0e})")
    assert e_gpu <= max(3.0 * e_cpu, tol)


def test_forced_low_rank_falls_back_to_adaptive():
    """Predicted rank 1 for every box: all sketches fail and the adaptive IDs are used."""
    _gpu()

    class _Low:
        def predict_rank(self, F):
            return np.ones(F.shape[0], dtype=np.int64)

    x = np.random.default_rng(0).random((30000, 3))
    f_a = _build(x, ops.LAPLACE_3D, 0.0, 1e-6, 128, "gpu")
    # a box of at most 10 candidates would be sketched in full (ceil(1.25) + 8 rows) and accepted
    assert f_a.bs_list[f_a.rank_list > 0].min() > 10
    f_l = _build(x, ops.LAPLACE_3D, 0.0, 1e-6, 128, "gpu", rank_model=_Low())
    st = f_l.gpu_stats
    assert st["sketched"] == 0 and st["direct"] == 0 and st["fallback"] > 0 and st["fallback"] == st["adaptive"]
    np.testing.assert_array_equal(f_l.rank_list, f_a.rank_list)
    for b in np.nonzero(f_a.rank_list)[0]:
        np.testing.assert_array_equal(f_l.cand[b], f_a.cand[b])
        np.testing.assert_array_equal(f_l.proj[b], f_a.proj[b])


# -------------------------------------------------------------------- MLP
def _mlp_dev(p):
    return torch.as_tensor(np.asarray(p, dtype=np.float32), device=_gpu()).contiguous()


def test_mlp_forward_matches_oracle():
    dev = _gpu()
    rng = np.random.default_rng(0)
    p = (rng.normal(size=mlp.NPARAM) * 0.3).astype(np.float32)
    X = rng.normal(size=(1000, 8)).astype(np.float32)
    m = lr.RankMLP(params=p, device=dev)
    y = m.forward_normalised(torch.as_tensor(X, device=dev)).cpu().numpy()
    ref = mlp.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.max(np.abs(y - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [1, 31, 32, 1000])
def test_mlp_backward_matches_oracle(n):
    dev = _gpu()
    L = hip.lib()
    rng = np.random.default_rng(n)
    p = (rng.normal(size=mlp.NPARAM) * 0.3).astype(np.float32)
    X = rng.normal(size=(n, 8)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    P, Xd, Td = _mlp_dev(p), _mlp_dev(X), _mlp_dev(t)
    g = torch.zeros(mlp.NPARAM, dtype=torch.float32, device=dev)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    hip.check_build(L.skf_mlp_backward(hip.ptr(P), hip.ptr(Xd), hip.ptr(Td), n, hip.ptr(g), hip.ptr(loss),
                                       hip.stream_handle(dev)))
    torch.cuda.synchronize()
    ref_loss, ref_g = mlp.loss_and_grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64))
    assert float(loss.item()) == pytest.approx(ref_loss, rel=1e-5)
    gh = g.cpu().numpy()
    # FP32 kernels with atomic accumulation: per-block bound relative to the block's largest entry
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mlp.SIZES])
    for a, b in zip(offs[:-1], offs[1:]):
        scale = max(np.max(np.abs(ref_g[a:b])), 1e-6)
        assert np.max(np.abs(gh[a:b] - ref_g[a:b])) <= 2e-5 * scale


def test_adam_kernel_matches_oracle():
    dev = _gpu()
    L = hip.lib()
    rng = np.random.default_rng(4)
    p = rng.normal(size=777).astype(np.float32)
    P = _mlp_dev(p)
    m1 = torch.zeros(777, dtype=torch.float32, device=dev)
    m2 = torch.zeros_like(m1)
    pr, a, b = p.astype(np.float64), np.zeros(777), np.zeros(777)
    for step in range(1, 6):
        g = rng.normal(size=777).astype(np.float32)
        gd = _mlp_dev(g)
        hip.check_build(L.skf_adam_step(hip.ptr(P), hip.ptr(gd), hip.ptr(m1), hip.ptr(m2), 777,
                                        ctypes.c_float(1e-3), ctypes.c_float(0.9), ctypes.c_float(0.999),
                                        ctypes.c_float(1e-8), step, hip.stream_handle(dev)))
        pr, a, b = mlp.adam(pr, g.astype(np.float64), a, b, step, 1e-3)
    torch.cuda.synchronize()
    np.testing.assert_allclose(P.cpu().numpy(), pr, rtol=0, atol=2e-6)


def _pairs():
    with np.load(lr.default_pairs_path(), allow_pickle=False) as z:
        return z["features"].astype(np.float64), z["ranks"].astype(np.int64)


def _group_mean_predictor(Ftr, rtr, Fte):
    """Baseline without the network: mean log rank of the training pairs with the same
    (operator, tolerance, level), else of the same (operator, tolerance), else overall."""
    y = np.log1p(rtr)
    key3 = lambda F: [tuple(v) for v in np.round(F[:, [7, 6, 0]]).astype(int)]
    key2 = lambda F: [tuple(v) for v in np.round(F[:, [7, 6]]).astype(int)]
    d3, d2 = {}, {}
    for k3, k2, v in zip(key3(Ftr), key2(Ftr), y):
        d3.setdefault(k3, []).append(v)
        d2.setdefault(k2, []).append(v)
    g = y.mean()
    out = []
    for k3, k2 in zip(key3(Fte), key2(Fte)):
        out.append(np.mean(d3[k3]) if k3 in d3 else (np.mean(d2[k2]) if k2 in d2 else g))
    return np.array(out)


def test_mlp_training_on_gpu_beats_rank_baseline():
    """Train from scratch with the HIP kernels; held-out log-rank error well below the
    baseline predictor's, and far fewer boxes predicted more than 20% too low."""
    dev = _gpu()
    F, r = _pairs()
    rng = np.random.default_rng(0)
    perm = rng.permutation(F.shape[0])
    nte = F.shape[0] // 5
    te, tr = perm[:nte], perm[nte:]
    m = lr.RankMLP(seed=1, device=dev).fit(F[tr], r[tr], epochs=25)
    assert m.history[-1] < 0.5 * m.history[0]
    pred = m.predict_log_rank(F[te])
    base = _group_mean_predictor(F[tr], r[tr], F[te])
    y = np.log1p(r[te])
    rmse_mlp = float(np.sqrt(np.mean((pred - y) ** 2)))
    rmse_base = float(np.sqrt(np.mean((base - y) ** 2)))
    low_mlp = float(np.mean(np.expm1(pred) * 1.25 + 8 < r[te]))
    low_base = float(np.mean(np.expm1(base) * 1.25 + 8 < r[te]))
    print(f"held-out log-rank rmse: mlp {rmse_mlp:.3f}  baseline {rmse_base:.3f}; "
          f"under-predicted sketches: mlp {low_mlp:.3%} baseline {low_base:.3%}")
    assert rmse_mlp < 0.6 * rmse_base
    assert low_mlp < 0.5 * low_base


def test_learned_build_accuracy_and_fallback_rate():
    _gpu()
    model = lr.load_default()
    assert model is not None
    # clustered cloud (16-component Gaussian mixture, sigma 0.02): many compressible boxes
    rng = np.random.default_rng(3)
    c = rng.random((16, 3))
    x = np.clip(c[rng.choice(16, 100000, p=rng.dirichlet(np.ones(16)))] + 0.02 * rng.normal(size=(100000, 3)), 0, 1)
    f = _build(x, ops.HELMHOLTZ_3D, 20.0, 1e-6, 256, "gpu", rank_model=model)
    st = f.gpu_stats
    rate = st["fallback"] / max(1, st["sketched"] + st["fallback"])
    e = _apply_err(f, x)
    print(f"learned build: {st}, fallback rate {rate:.2%}, apply error {e:.2e}")
    assert st["sketched"] > 0 and rate < 0.15
    assert st["adaptive"] == st["fallback"] + st["direct"]
    assert e < 1e-5
