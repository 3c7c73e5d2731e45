"""Learned skeleton-rank predictor for the GPU build.

The reference finds every box's skeleton rank by a full pivoted QR of its
proxy matrix followed by a threshold count (``skelfmm/util.py:13-51``). Here an
8 -> 64 -> 64 -> 1 GELU MLP in FP32 maps box features to the rank so that the
GPU build can factor a Gaussian sketch of ``ceil(1.25 k) + 8`` rows instead of
the whole proxy matrix (``skelfmm_amd.gpu_build.GPUSkeletonizer.sketched``).

Features of a box (``box_features``), in this order:

0. tree level;
1. log of the candidate count;
2-4. eigenvalues of the candidates' second-moment matrix about their mean,
     divided by the squared box side, largest first (zero third one in 2D);
5. wavenumber times box side;
6. log10 of the ID tolerance;
7. operator code (``KernelOperator.kernel_id``).

The network predicts ``log(1 + rank)`` after standardisation of inputs and
target. Training pairs (box features, adaptive-ID rank) are harvested from host
builds (``harvest``); training runs on the GPU with the library's forward,
backward and Adam kernels (``skf_mlp_forward``, ``skf_mlp_backward``,
``skf_adam_step``).
"""

from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import torch

from . import hip
from . import operators as operators_mod

__all__ = ["NFEAT", "NHID", "box_features", "RankMLP", "harvest", "harvest_suite", "DATA_DIR",
           "default_model_path", "default_pairs_path", "load_default"]

NFEAT = 8
NHID = 64
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def default_pairs_path():
    return os.path.join(DATA_DIR, "rank_pairs.npz")


def default_model_path():
    return os.path.join(DATA_DIR, "rank_mlp.npz")


def param_count():
    return NHID * NFEAT + NHID + NHID * NHID + NHID + NHID + 1


# ----------------------------------------------------------------- features
def box_features(skel, lev, boxes, centers=None, length=None):
    """(len(boxes), 8) float64 features of the boxes of one level of a build."""
    boxes = np.asarray(boxes, dtype=np.int64)
    nb = boxes.shape[0]
    F = np.zeros((nb, NFEAT))
    if nb == 0:
        return F
    if length is None:
        length = skel.tree.get_box_length(int(boxes[0]))
    d = skel.ndim
    cands = [skel.cand[b] for b in boxes]
    cnt = np.array([c.shape[0] for c in cands], dtype=np.int64)
    X = skel.tree.XX[np.concatenate(cands)] if cnt.sum() else np.zeros((0, d))
    starts = np.concatenate(([0], np.cumsum(cnt)[:-1]))
    safe = np.maximum(cnt, 1)[:, None]
    nz = cnt > 0
    mean = np.zeros((nb, d))
    M = np.zeros((nb, 3, 3))
    if X.shape[0]:
        s1 = np.zeros((nb, d))
        s1[nz] = np.add.reduceat(X, starts[nz], axis=0)
        mean = s1 / safe
        for a in range(d):
            for b in range(a, d):
                sab = np.zeros(nb)
                sab[nz] = np.add.reduceat(X[:, a] * X[:, b], starts[nz])
                M[:, a, b] = M[:, b, a] = sab / safe[:, 0] - mean[:, a] * mean[:, b]
    ev = np.linalg.eigvalsh(M)[:, ::-1] / (length * length)
    F[:, 0] = lev
    F[:, 1] = np.log(np.maximum(cnt, 1))
    F[:, 2:5] = np.maximum(ev, 0.0)
    F[:, 5] = skel.kappa * length
    F[:, 6] = math.log10(skel.tol)
    F[:, 7] = skel.operator.kernel_id
    return F


# --------------------------------------------------------------------- model
def _init_params(seed):
    """He-style initialisation of W1, b1, W2, b2, w3, b3 (FP32, flat)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, math.sqrt(2.0 / NFEAT), (NHID, NFEAT))
    W2 = rng.normal(0, math.sqrt(2.0 / NHID), (NHID, NHID))
    w3 = rng.normal(0, math.sqrt(1.0 / NHID), NHID)
    return np.concatenate([W1.ravel(), np.zeros(NHID), W2.ravel(), np.zeros(NHID), w3, [0.0]]).astype(np.float32)


class RankMLP:
    """The rank network: normalisation constants plus the flat FP32 parameter vector."""

    def __init__(self, params=None, x_mean=None, x_std=None, y_mean=0.0, y_std=1.0, seed=0, device=None):
        self.params_host = _init_params(seed) if params is None else np.asarray(params, dtype=np.float32).copy()
        if self.params_host.shape != (param_count(),):
            raise ValueError(f"rank MLP expects {param_count()} parameters, got {self.params_host.shape}")
        self.x_mean = np.zeros(NFEAT) if x_mean is None else np.asarray(x_mean, dtype=np.float64)
        self.x_std = np.ones(NFEAT) if x_std is None else np.asarray(x_std, dtype=np.float64)
        self.y_mean = float(y_mean)
        self.y_std = float(y_std)
        self.device = device
        self._params_dev = None
        self.history = []

    # ------------------------------------------------------------- device
    def _dev(self):
        hip.require_gpu()
        if self.device is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        if self._params_dev is None:
            self._params_dev = torch.as_tensor(self.params_host, device=self.device).contiguous()
        return self._params_dev

    def _normalise(self, F):
        Xn = (np.asarray(F, dtype=np.float64) - self.x_mean) / self.x_std
        return torch.as_tensor(Xn.astype(np.float32), device=self.device).contiguous()

    # ---------------------------------------------------------------- use
    def forward_normalised(self, X):
        """Network output for already-standardised FP32 inputs on the device."""
        P = self._dev()
        L = hip.lib()
        n = int(X.shape[0])
        y = torch.empty(max(n, 1), dtype=torch.float32, device=self.device)
        if n:
            hip.check_build(L.skf_mlp_forward(hip.ptr(P), hip.ptr(X), n, hip.ptr(y), hip.stream_handle(self.device)))
        return y[:n]

    def predict_log_rank(self, F):
        self._dev()
        y = self.forward_normalised(self._normalise(F))
        return y.double().cpu().numpy() * self.y_std + self.y_mean

    def predict_rank(self, F):
        """Predicted skeleton rank per box (>= 1, at most the candidate count)."""
        F = np.asarray(F)
        if F.shape[0] == 0:
            return np.zeros(0, dtype=np.int64)
        k = np.expm1(self.predict_log_rank(F))
        bs = np.exp(F[:, 1])
        return np.clip(np.rint(k), 1, np.maximum(np.rint(bs), 1)).astype(np.int64)

    # -------------------------------------------------------------- train
    def fit(self, F, ranks, *, epochs=60, batch=4096, lr=3e-3, lr_final=3e-4, seed=0, verbose=False):
        """Adam on the mean squared error of the standardised log rank (GPU kernels)."""
        F = np.asarray(F, dtype=np.float64)
        y = np.log1p(np.asarray(ranks, dtype=np.float64))
        self.x_mean = F.mean(axis=0)
        self.x_std = np.where(F.std(axis=0) > 1e-12, F.std(axis=0), 1.0)
        self.y_mean = float(y.mean())
        self.y_std = float(y.std()) if y.std() > 1e-12 else 1.0
        self._dev()
        L = hip.lib()
        st = hip.stream_handle(self.device)
        X = self._normalise(F)
        Y = torch.as_tensor(((y - self.y_mean) / self.y_std).astype(np.float32), device=self.device)
        P = self._params_dev
        g = torch.zeros_like(P)
        m1 = torch.zeros_like(P)
        m2 = torch.zeros_like(P)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        n = X.shape[0]
        gen = torch.Generator(device="cpu")
        gen.manual_seed(seed)
        step = 0
        nsteps = epochs * max(1, (n + batch - 1) // batch)
        self.history = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).to(self.device)
            Xs, Ys = X[perm].contiguous(), Y[perm].contiguous()
            tot = torch.zeros(1, dtype=torch.float32, device=self.device)
            for s0 in range(0, n, batch):
                nb = min(batch, n - s0)
                g.zero_()
                loss.zero_()
                hip.check_build(L.skf_mlp_backward(hip.ptr(P), hip.ptr(Xs[s0:]), hip.ptr(Ys[s0:]), nb, hip.ptr(g),
                                                   hip.ptr(loss), st))
                step += 1
                frac = step / nsteps
                lr_t = lr * (lr_final / lr) ** frac  # geometric decay
                hip.check_build(L.skf_adam_step(hip.ptr(P), hip.ptr(g), hip.ptr(m1), hip.ptr(m2), P.numel(),
                                                ctypes.c_float(lr_t), ctypes.c_float(0.9), ctypes.c_float(0.999),
                                                ctypes.c_float(1e-8), step, st))
                tot += loss * nb
            self.history.append(float(tot.item()) / n)
            if verbose:
                print(f"epoch {ep:3d}  mse {self.history[-1]:.5f}")
        self.params_host = P.cpu().numpy().copy()
        return self

    # ----------------------------------------------------------------- io
    def save(self, path):
        np.savez(path, params=self.params_host, x_mean=self.x_mean, x_std=self.x_std,
                 y_mean=self.y_mean, y_std=self.y_std)

    @classmethod
    def load(cls, path, device=None):
        with np.load(path, allow_pickle=False) as z:
            return cls(params=z["params"], x_mean=z["x_mean"], x_std=z["x_std"], y_mean=float(z["y_mean"]),
                       y_std=float(z["y_std"]), device=device)


def load_default(device=None):
    """The committed trained network, or None if it is absent."""
    p = default_model_path()
    return RankMLP.load(p, device=device) if os.path.exists(p) else None


# ------------------------------------------------------------------- harvest
def harvest(points, operator, kappa, tol, leaf_size, geometry=None, npoints_max=500):
    """(features, adaptive rank) of every skeletonized box of one host build."""
    from .build import RecursiveSkeletonization
    from .tree import BalancedTree

    tree = BalancedTree(points, leaf_size)
    skel = RecursiveSkeletonization(tree, operator, kappa=kappa, tol=tol, geometry=geometry)
    skel.skel_tree(npoints_max=npoints_max)
    Fs, rs = [], []
    for lev in range(skel.root_level + 1, tree.nlevels):
        boxes = tree.get_boxes_level(lev)
        boxes = boxes[skel.bs_list[boxes] > 0]
        if boxes.shape[0] == 0:
            continue
        Fs.append(box_features(skel, lev, boxes))
        rs.append(skel.rank_list[boxes])
    if not Fs:
        return np.zeros((0, NFEAT)), np.zeros(0, dtype=np.int64)
    return np.concatenate(Fs), np.concatenate(rs).astype(np.int64)


def _cloud(kind, n, dim, rng, ncomp=16, sigma=0.02):
    if kind == "cube":
        return rng.random((n, dim)), None
    if kind == "sphere":
        v = rng.normal(size=(n, 3))
        x = v / np.linalg.norm(v, axis=1, keepdims=True)
        return x, x.copy()
    if kind == "mixture":
        c = rng.random((ncomp, dim))
        w = rng.dirichlet(np.ones(ncomp))
        comp = rng.choice(ncomp, size=n, p=w)
        x = np.clip(c[comp] + sigma * rng.normal(size=(n, dim)), 0.0, 1.0)
        return x, None
    if kind == "annulus":
        th = rng.random(n) * 2 * np.pi
        r = (1 + 0.25 * np.cos(6 * th)) * (0.5 + 0.5 * rng.random(n))
        return np.stack([r * np.cos(th), r * np.sin(th)], axis=1), None
    raise ValueError(kind)


def harvest_suite(n_builds=48, seed=0, max_points=60000, verbose=False, suite="mixed"):
    """Pairs from host builds over uniform-cube, sphere-surface and mixture clouds.

    ``suite="mixed"``: operators, wavenumbers in [0, 40], tolerances in {1e-4, ..., 1e-8}
    and leaf sizes drawn per build from ``default_rng(seed)``. ``suite="clustered"``:
    large 3D Gaussian mixtures (8-32 components, sigma 0.01-0.05), Helmholtz or Laplace,
    tol 1e-5..1e-7, leaf 128 or 256 -- deep adaptive trees.

    The committed set ``data/rank_pairs.npz`` (383,128 pairs from 119 builds) is
    ``harvest --builds 72 --seed 0 --max-points 80000``, then ``--builds 32 --seed 1
    --max-points 80000 --append``, then ``--suite clustered --builds 16 --seed 2
    --max-points 600000 --append`` (one of those 16 builds is skipped: clipped points pile up
    on a cube corner beyond the leaf size).
    """
    O = operators_mod
    if suite == "clustered":
        return _harvest_clustered(n_builds, seed, max_points, verbose)
    choices_3d = [O.LAPLACE_3D, O.HELMHOLTZ_3D, O.HELMHOLTZ_3D, O.LAPLACE_DOUBLE_LAYER_3D,
                  O.HELMHOLTZ_DOUBLE_LAYER_3D, O.LAPLACE_ADJOINT_DOUBLE_LAYER_3D,
                  O.HELMHOLTZ_ADJOINT_DOUBLE_LAYER_3D]
    rng = np.random.default_rng(seed)
    Fs, rs, meta = [], [], []
    for b in range(n_builds):
        two_d = rng.random() < 0.2
        if two_d:
            op = O.LAPLACE_2D if rng.random() < 0.6 else O.HELMHOLTZ_2D
            kind = ["cube", "mixture", "annulus"][rng.integers(3)]
            dim = 2
        else:
            op = choices_3d[rng.integers(len(choices_3d))]
            kind = ["cube", "sphere", "mixture"][rng.integers(3)]
            if op.requires_geometry:
                kind = "sphere"
            dim = 3
        kappa = float(rng.uniform(0, 40)) if "HELMHOLTZ" in op.name.upper() else 0.0
        tol = float(10.0 ** -rng.integers(4, 9))
        leaf = int(rng.choice([64, 96, 128, 192, 256]))
        n = int(rng.integers(max_points // 4, max_points + 1))
        x, normals = _cloud(kind, n, dim, rng)
        geom = normals if op.requires_geometry else None
        F, r = harvest(x, op, kappa, tol, leaf, geometry=geom)
        Fs.append(F)
        rs.append(r)
        meta.append((op.name, kind, n, kappa, tol, leaf, F.shape[0]))
        if verbose:
            print(f"build {b:3d}: {op.name:32s} {kind:8s} N={n:6d} kappa={kappa:5.1f} tol={tol:.0e} "
                  f"leaf={leaf:3d} -> {F.shape[0]} pairs", flush=True)
    return np.concatenate(Fs), np.concatenate(rs), meta


def _harvest_clustered(n_builds, seed, max_points, verbose):
    O = operators_mod
    rng = np.random.default_rng(seed)
    Fs, rs, meta = [], [], []
    for b in range(n_builds):
        op = O.HELMHOLTZ_3D if rng.random() < 0.75 else O.LAPLACE_3D
        kappa = float(rng.uniform(0, 40)) if op is O.HELMHOLTZ_3D else 0.0
        tol = float(10.0 ** -rng.integers(5, 8))
        leaf = int(rng.choice([128, 256]))
        n = int(rng.integers(max_points // 4, max_points + 1))
        ncomp = int(rng.choice([8, 16, 32]))
        sigma = float(rng.choice([0.01, 0.02, 0.05]))
        x, _ = _cloud("mixture", n, 3, rng, ncomp=ncomp, sigma=sigma)
        try:
            F, r = harvest(x, op, kappa, tol, leaf)
        except ValueError as e:  # e.g. clipped points piling up on a cube corner beyond leaf_size
            if verbose:
                print(f"build {b:3d}: skipped ({e})", flush=True)
            continue
        Fs.append(F)
        rs.append(r)
        meta.append((op.name, f"mixture{ncomp}/{sigma}", n, kappa, tol, leaf, F.shape[0]))
        if verbose:
            print(f"build {b:3d}: {op.name:14s} mixture {ncomp:2d}x{sigma:.2f} N={n:7d} kappa={kappa:5.1f} "
                  f"tol={tol:.0e} leaf={leaf:3d} -> {F.shape[0]} pairs", flush=True)
    return np.concatenate(Fs), np.concatenate(rs), meta


def _main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(description="harvest (box features, adaptive rank) pairs / train the rank MLP")
    sub = ap.add_subparsers(dest="cmd", required=True)
    h = sub.add_parser("harvest")
    h.add_argument("--builds", type=int, default=48)
    h.add_argument("--seed", type=int, default=0)
    h.add_argument("--max-points", type=int, default=60000)
    h.add_argument("--out", default=default_pairs_path())
    h.add_argument("--append", action="store_true", help="add to the pairs already in --out")
    h.add_argument("--suite", choices=("mixed", "clustered"), default="mixed")
    t = sub.add_parser("train")
    t.add_argument("--pairs", default=default_pairs_path())
    t.add_argument("--out", default=default_model_path())
    t.add_argument("--epochs", type=int, default=60)
    a = ap.parse_args(argv)
    if a.cmd == "harvest":
        F, r, meta = harvest_suite(a.builds, a.seed, a.max_points, verbose=True, suite=a.suite)
        F, r = F.astype(np.float32), r.astype(np.int32)
        builds = np.array([" ".join(map(str, m)) for m in meta])
        if a.append and os.path.exists(a.out):
            with np.load(a.out, allow_pickle=False) as z:
                F = np.concatenate([z["features"], F])
                r = np.concatenate([z["ranks"], r])
                builds = np.concatenate([z["builds"], builds])
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        np.savez_compressed(a.out, features=F, ranks=r, builds=builds)
        print(f"{F.shape[0]} pairs -> {a.out}")
    else:
        with np.load(a.pairs, allow_pickle=False) as z:
            F, r = z["features"].astype(np.float64), z["ranks"]
        model = RankMLP().fit(F, r, epochs=a.epochs, verbose=True)
        model.save(a.out)
        print(f"trained on {F.shape[0]} pairs -> {a.out}")


if __name__ == "__main__":
    _main()
