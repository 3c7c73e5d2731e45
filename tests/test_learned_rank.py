"""CPU tests of the learned-rank build's host logic and of the MLP oracle.

* ``oracle.mlp``: hand-written backpropagation against central finite
  differences of the forward pass; Adam's first step against its closed form.
* ``learned_rank.box_features``: invariances and closed forms of the geometric
  features; harvested pairs equal the host build's ranks.
* ``gpu_build.accept_sketch``: the proxy-residual acceptance test on sketched
  IDs computed here with scipy -- a deliberately low predicted rank, or a
  corrupted interpolation matrix, is rejected (the fallback is forced).
* The committed training pairs and trained network are well formed, and the
  predictor fails loudly without a GPU.
"""

import os

import numpy as np
import pytest
import torch
from scipy.linalg import qr, solve_triangular

from conftest import make_build
from oracle import mlp
from skelfmm_amd import learned_rank as lr
from skelfmm_amd import operators as ops
from skelfmm_amd.gpu_build import accept_sketch


# ---------------------------------------------------------------- MLP oracle
def _fd_grad(p, X, t, idx, h=1e-6):
    out = []
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp = np.mean((mlp.forward(p + e, X) - t) ** 2)
        lm = np.mean((mlp.forward(p - e, X) - t) ** 2)
        out.append((lp - lm) / (2 * h))
    return np.array(out)


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_mlp_gradient_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=mlp.NPARAM) * 0.3
    X = rng.normal(size=(9, mlp.NIN))
    t = rng.normal(size=9)
    loss, g = mlp.loss_and_grad(p, X, t)
    assert loss == pytest.approx(np.mean((mlp.forward(p, X) - t) ** 2), rel=1e-14)
    # every parameter block: W1, b1, W2, b2, w3, b3
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mlp.SIZES])
    idx = np.concatenate([rng.choice(np.arange(a, b), min(8, b - a), replace=False) for a, b in zip(offs[:-1], offs[1:])])
    fd = _fd_grad(p, X, t, idx)
    assert np.max(np.abs(fd - g[idx])) <= 1e-7 * max(1.0, np.max(np.abs(g[idx])))


def test_oracle_gelu_is_z_phi():
    from scipy.stats import norm

    z = np.linspace(-6, 6, 101)
    np.testing.assert_allclose(mlp.gelu(z), z * norm.cdf(z), rtol=1e-13, atol=1e-15)
    h = 1e-6
    np.testing.assert_allclose(mlp.gelu_d(z), (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h), rtol=1e-6, atol=1e-9)


def test_oracle_adam_first_step_is_signed_lr():
    rng = np.random.default_rng(3)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, m1, m2 = mlp.adam(p, g, np.zeros(50), np.zeros(50), 1, lr=1e-2, eps=0.0)
    np.testing.assert_allclose(p1, p - 1e-2 * np.sign(g), rtol=0, atol=1e-15)
    np.testing.assert_allclose(m1, 0.1 * g)
    np.testing.assert_allclose(m2, 0.001 * g * g)


def test_rank_mlp_parameter_layout_matches_oracle():
    assert lr.param_count() == mlp.NPARAM == 4801


def test_planted_transposed_weight_fails_gradient_check():
    """A backprop that uses W2 instead of W2^T is caught by the finite-difference check."""
    rng = np.random.default_rng(5)
    p = rng.normal(size=mlp.NPARAM) * 0.3
    X = rng.normal(size=(6, mlp.NIN))
    t = rng.normal(size=6)
    P = mlp.unpack(p)
    z1 = X @ P["W1"].T + P["b1"]
    z2 = mlp.gelu(z1) @ P["W2"].T + P["b2"]
    y = mlp.gelu(z2) @ P["w3"] + P["b3"][0]
    d2 = np.outer(2 * (y - t) / 6, P["w3"]) * mlp.gelu_d(z2)
    d1_bad = (d2 @ P["W2"].T) * mlp.gelu_d(z1)
    g_b1_bad = d1_bad.sum(axis=0)
    off = mlp.NH * mlp.NIN
    fd = _fd_grad(p, X, t, np.arange(off, off + mlp.NH))
    assert np.max(np.abs(fd - g_b1_bad)) > 1e-3 * np.max(np.abs(fd))


# ------------------------------------------------------------------ features
class _FakeTree:
    def __init__(self, X, L):
        self.XX = X
        self._L = L

    def get_box_length(self, b):
        return self._L


class _FakeSkel:
    def __init__(self, X, cands, L, kappa=2.0, tol=1e-6, op=ops.HELMHOLTZ_3D):
        self.tree = _FakeTree(X, L)
        self.ndim = X.shape[1]
        self.cand = cands
        self.kappa = kappa
        self.tol = tol
        self.operator = op


def test_box_features_closed_forms_and_invariance():
    rng = np.random.default_rng(0)
    n = 200000
    X = rng.random((n, 3))
    s = _FakeSkel(X, [np.arange(n), np.arange(10)], 1.0)
    F = lr.box_features(s, 4, [0, 1])
    assert F.shape == (2, 8)
    assert F[0, 0] == 4 and F[0, 1] == pytest.approx(np.log(n))
    np.testing.assert_allclose(F[0, 2:5], 1 / 12, rtol=2e-2)  # uniform cube: variance L^2/12 per axis
    assert np.all(np.diff(F[:, 2:5], axis=1) <= 0)
    assert F[0, 5] == pytest.approx(2.0) and F[0, 6] == pytest.approx(-6) and F[0, 7] == ops.HELMHOLTZ_3D.kernel_id
    # translation, rotation and scale (with the box side) leave the moment features unchanged
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    X2 = 3.7 * (X @ Q.T) + np.array([5.0, -2.0, 1.0])
    s2 = _FakeSkel(X2, s.cand, 3.7)
    F2 = lr.box_features(s2, 4, [0, 1])
    np.testing.assert_allclose(F2[:, 2:5], F[:, 2:5], rtol=1e-9, atol=1e-12)
    assert F2[0, 5] == pytest.approx(2.0 * 3.7)


def test_box_features_planar_and_2d():
    rng = np.random.default_rng(1)
    X = np.zeros((5000, 3))
    X[:, :2] = rng.random((5000, 2))
    F = lr.box_features(_FakeSkel(X, [np.arange(5000)], 1.0), 2, [0])
    assert F[0, 4] == pytest.approx(0.0, abs=1e-15)
    X2 = rng.random((3000, 2))
    F2 = lr.box_features(_FakeSkel(X2, [np.arange(3000)], 1.0, op=ops.LAPLACE_2D), 3, [0])
    np.testing.assert_allclose(F2[0, 2:4], 1 / 12, rtol=5e-2)
    assert F2[0, 4] == 0.0 and F2[0, 7] == 0


def test_harvest_pairs_are_the_builds_ranks():
    rng = np.random.default_rng(0)
    x = rng.random((6000, 3))
    F, r = lr.harvest(x, ops.LAPLACE_3D, 0.0, 1e-5, 64)
    fmm = make_build(x, ops.LAPLACE_3D, tol=1e-5, leaf_size=64)
    tree = fmm.tree
    want = []
    for lev in range(fmm.root_level + 1, tree.nlevels):
        b = tree.get_boxes_level(lev)
        want.append(fmm.rank_list[b[fmm.bs_list[b] > 0]])
    np.testing.assert_array_equal(r, np.concatenate(want))
    assert F.shape == (r.shape[0], 8) and np.all(r >= 1)
    assert np.all(np.exp(F[:, 1]) + 0.5 >= r)  # rank never exceeds the candidate count


# ---------------------------------------------------------- sketch acceptance
def _sketch_id(A, l, tol, rng):
    """Sketched ID computed here with scipy: QR of (Omega A)[:l] stopped at the cutoff."""
    Om = rng.normal(size=(l + 8, A.shape[0]))
    YZ = Om @ A
    Y, Z = YZ[:l], YZ[l:]
    R, piv = qr(Y, mode="r", pivoting=True)
    d = np.abs(np.diag(R))
    rank = int(np.sum(d > tol * d[0]))
    T = solve_triangular(R[:rank, :rank], R[:rank, rank:])
    return (rank, piv, T), Z


def _lowrank(m, n, k, rng, decay=1e-12):
    U, _ = np.linalg.qr(rng.normal(size=(m, min(m, n))))
    V, _ = np.linalg.qr(rng.normal(size=(n, min(m, n))))
    s = np.concatenate([np.logspace(0, -4, k), np.full(min(m, n) - k, decay)])
    return (U * s) @ V.T


def test_accept_sketch_accepts_sufficient_rank():
    rng = np.random.default_rng(0)
    A = _lowrank(300, 120, 30, rng)
    res, Z = _sketch_id(A, 46, 1e-8, rng)
    assert res[0] < 46
    assert accept_sketch(Z, res, 46, 1e-8, cap=120)


def test_accept_sketch_rejects_low_predicted_rank():
    """The forced fallback: a sketch smaller than the true rank is rejected."""
    rng = np.random.default_rng(1)
    A = _lowrank(300, 120, 30, rng)
    res, Z = _sketch_id(A, 12, 1e-8, rng)
    assert res[0] == 12
    assert not accept_sketch(Z, res, 12, 1e-8, cap=120)


def test_accept_sketch_rejects_bad_interpolation():
    rng = np.random.default_rng(2)
    A = _lowrank(300, 120, 30, rng)
    res, Z = _sketch_id(A, 46, 1e-8, rng)
    rank, piv, T = res
    T_bad = T.copy()
    T_bad[0, :] *= 1.0 + 1e-5
    assert not accept_sketch(Z, (rank, piv, T_bad), 46, 1e-8, cap=120)


def test_accept_sketch_full_rank_box_at_cap():
    rng = np.random.default_rng(3)
    A = rng.normal(size=(200, 20))
    res, Z = _sketch_id(A, 20, 1e-8, rng)
    assert res[0] == 20
    assert accept_sketch(Z, res, 20, 1e-8, cap=20)
    assert not accept_sketch(Z, res, 20, 1e-8, cap=200)


# ------------------------------------------------------------ data and model
def test_committed_training_pairs_are_well_formed():
    path = lr.default_pairs_path()
    assert os.path.exists(path), "training pairs missing"
    with np.load(path, allow_pickle=False) as z:
        F, r = z["features"], z["ranks"]
    assert F.shape[1] == 8 and F.shape[0] == r.shape[0] >= 50000
    assert np.all(r >= 1) and np.all(np.exp(F[:, 1]) + 0.5 >= r)
    tols = np.unique(np.round(F[:, 6]))
    assert set(tols.tolist()) == {-8.0, -7.0, -6.0, -5.0, -4.0}
    assert F[:, 5].max() > 1.0 and set(np.unique(F[:, 7]).astype(int).tolist()) >= {0, 1, 3}


def test_committed_model_roundtrip(tmp_path):
    m = lr.load_default()
    assert m is not None, "trained rank network missing"
    assert m.params_host.shape == (4801,) and np.all(np.isfinite(m.params_host))
    p = tmp_path / "m.npz"
    m.save(p)
    m2 = lr.RankMLP.load(p)
    np.testing.assert_array_equal(m2.params_host, m.params_host)
    np.testing.assert_array_equal(m2.x_std, m.x_std)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_predictor_fails_loudly_without_gpu():
    m = lr.RankMLP(seed=0)
    with pytest.raises(RuntimeError):
        m.predict_rank(np.zeros((3, 8)))
