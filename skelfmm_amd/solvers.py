"""Krylov solvers that use the fast apply as their operator (second-kind integral equations).

* ``FMMOperator(fmm, weights=None, diag=None, graph=False)``: ``A x = diag * x + K~ (weights * x)``
  on the device, ``K~`` the skeletonized apply of a ``SkelFMM`` or ``SkelFMMBatched``
  (quadrature weights scale the charges; ``diag`` carries the identity term of a
  second-kind equation and any local singular correction). With ``graph=True`` the apply
  runs as one HIP graph replay per product on fixed device buffers.
* ``gmres(A, b, ...)``: restarted GMRES(m) with classical Gram-Schmidt applied twice
  (CGS2). Krylov vectors stay on the device; the small Hessenberg matrix and its Givens
  rotations live on the host. The columns of a block ``b`` (``(N, R)``) iterate in
  lockstep and share every apply (one multi-RHS apply per step), each with its own
  Krylov space, rotations and stopping test.
* ``laplace_sphere_dirichlet(n, ...)``: the interior Dirichlet problem on the unit sphere
  as the double-layer equation ``(1/2 + c_j) sigma_j + sum_k K(x_j, x_k) w_k sigma_k = f_j``
  -- the kind of boundary-integral solve the reference's sphere studies drive
  (``skelfmm/examples/util/sphere_kernel_utils.py:167-257``, GMRES iteration table
  ``skelfmm/examples/util/report_utils.py:62-95``).

The double-layer kernel of this library is ``K(x, y) = -(x - y).n_y / (4 pi r^3)``; with
the outward normal it integrates to 1 inside the surface, 1/2 on it and 0 outside, so the
interior limit of ``u = int K sigma`` is ``sigma / 2 + int K sigma``. On the unit sphere
``K(x, y) = 1 / (8 pi |x - y|)``; the point rule leaves out the self cell, whose integral
over a flat disc of the cell's area ``w_j`` is ``c_j = sqrt(w_j / pi) / 4`` (the local
correction ``correction=True`` adds to the diagonal).
"""

from __future__ import annotations

import math
from time import perf_counter

import numpy as np
import torch
from scipy.linalg import solve_triangular

__all__ = ["FMMOperator", "gmres", "GMRESResult", "fibonacci_sphere", "laplace_sphere_dirichlet"]


class FMMOperator:
    """``x -> diag * x + K~ (weights * x)`` for device tensors of shape ``(N,)`` or ``(N, R)``."""

    def __init__(self, fmm, weights=None, diag=None, graph=False):
        if hasattr(fmm, "apply_device"):
            dev = fmm
        else:
            fmm._ensure_built()
            dev = fmm.batched
        self.dev = dev
        self.N = int(dev.N)
        self.device = dev.device
        self.dtype = dev.complex_dtype if dev.complex_vectors else dev.real_dtype
        self.weights = None if weights is None else torch.as_tensor(weights, dtype=dev.real_dtype,
                                                                    device=self.device)
        if diag is None or np.isscalar(diag):
            self.diag = diag
        else:
            self.diag = torch.as_tensor(diag, device=self.device).to(self.dtype)
        self.graph = bool(graph)
        self._graphs = {}
        self.napplies = 0

    def _kernel_apply(self, q):
        if not self.graph:
            return self.dev.apply_device(q)
        key = tuple(q.shape)
        held = self._graphs.get(key)
        if held is None or held[2].stale:  # first use, or the apply's settings changed since
            qb = torch.empty_like(q)
            yb = torch.empty_like(q)
            qb.copy_(q)
            self._graphs[key] = (qb, yb, self.dev.capture(qb, yb))
        qb, yb, g = self._graphs[key]
        qb.copy_(q)
        g.replay()
        return yb.clone()

    def __call__(self, x):
        if x.shape[0] != self.N:
            raise ValueError("vector length must match the number of points")
        x = x.to(self.dtype)
        q = x
        if self.weights is not None:
            q = x * (self.weights if x.dim() == 1 else self.weights[:, None])
        y = self._kernel_apply(q.contiguous())
        self.napplies += 1 if x.dim() == 1 else x.shape[1]
        if self.diag is not None:
            if torch.is_tensor(self.diag):
                y = y + (self.diag if x.dim() == 1 else self.diag[:, None]) * x
            else:
                y = y + self.diag * x
        return y


class GMRESResult:
    """Solution and history of ``gmres``: ``x``, per-column ``iterations``, ``converged``,
    ``residuals`` (relative residual norms per step, per column) and ``restarts``."""

    def __init__(self, x, iterations, converged, residuals, restarts):
        self.x = x
        self.iterations = iterations
        self.converged = converged
        self.residuals = residuals
        self.restarts = restarts

    def __repr__(self):
        return f"GMRESResult(iterations={self.iterations.tolist()}, converged={self.converged.tolist()})"


def _norms(V):
    """Column 2-norms of a (N, R) tensor in float64 on the host."""
    return torch.linalg.vector_norm(V, dim=0).to(torch.float64).cpu().numpy()


def _givens(a, b):
    """Rotation (c real, s) with ``[[c, s], [-conj(s), c]] @ [a, b] = [rho, 0]`` (``b`` real >= 0)."""
    absa = np.abs(a)
    r = np.hypot(absa, b)
    c = np.where(r > 0, absa / np.where(r > 0, r, 1.0), 1.0)
    phase = np.where(absa > 0, a / np.where(absa > 0, absa, 1.0), 1.0)
    s = np.where(r > 0, phase * b / np.where(r > 0, r, 1.0), 0.0)
    return c, s


def gmres(A, b, x0=None, *, tol=1e-8, restart=30, maxiter=300, callback=None):
    """Solve ``A x = b`` by restarted GMRES; ``b`` is ``(N,)`` or ``(N, R)`` (columns solved together).

    ``A`` maps a device tensor of ``b``'s shape to one of the same shape. A column stops
    when its relative residual ``||b - A x|| / ||b||`` is at most ``tol`` (a zero column is
    solved by zero); ``maxiter`` counts Krylov steps per column over all restarts.
    """
    one = b.dim() == 1
    B = b[:, None] if one else b
    N, R = B.shape
    cplx = B.is_complex()
    hdt = np.complex128 if cplx else np.float64
    bnorm = _norms(B)
    X = torch.zeros_like(B) if x0 is None else (x0[:, None] if one else x0).clone().to(B.dtype)
    iters = np.zeros(R, dtype=np.int64)
    converged = bnorm == 0
    history = [[] for _ in range(R)]
    restarts = 0
    m = max(1, int(restart))

    def Aop(V):
        return A(V[:, 0])[:, None] if one else A(V)

    while True:
        Rv = B - Aop(X) if (x0 is not None or restarts > 0) else B.clone()
        beta = _norms(Rv)
        rel = np.where(bnorm > 0, beta / np.where(bnorm > 0, bnorm, 1.0), 0.0)
        converged = converged | (rel <= tol)
        if converged.all() or (iters >= maxiter).all():
            break
        active = ~converged & (iters < maxiter)
        Vs = torch.zeros((m + 1, N, R), dtype=B.dtype, device=B.device)
        inv = torch.as_tensor(np.where(beta > 0, 1.0 / np.where(beta > 0, beta, 1.0), 0.0), device=B.device)
        Vs[0] = Rv * inv.to(Rv.real.dtype if cplx else Rv.dtype)
        H = np.zeros((R, m + 1, m), dtype=hdt)
        cs = np.zeros((R, m))
        sn = np.zeros((R, m), dtype=hdt)
        g = np.zeros((R, m + 1), dtype=hdt)
        g[:, 0] = beta
        steps = np.zeros(R, dtype=np.int64)  # Krylov steps this cycle that count for each column
        live = active.copy()
        for j in range(m):
            W = Aop(Vs[j])
            basis = Vs[:j + 1]
            h = torch.einsum("knr,nr->kr", basis.conj(), W)
            W = W - torch.einsum("knr,kr->nr", basis, h)
            h2 = torch.einsum("knr,nr->kr", basis.conj(), W)
            W = W - torch.einsum("knr,kr->nr", basis, h2)
            hcol = (h + h2).to(torch.complex128 if cplx else torch.float64).cpu().numpy()
            hn = _norms(W)
            Vs[j + 1] = W * torch.as_tensor(np.where(hn > 0, 1.0 / np.where(hn > 0, hn, 1.0), 0.0),
                                            device=B.device).to(W.real.dtype if cplx else W.dtype)
            H[:, :j + 1, j] = hcol.T
            H[:, j + 1, j] = hn
            for i in range(j):  # earlier rotations on the new column
                hi, hi1 = H[:, i, j].copy(), H[:, i + 1, j].copy()
                H[:, i, j] = cs[:, i] * hi + sn[:, i] * hi1
                H[:, i + 1, j] = -np.conj(sn[:, i]) * hi + cs[:, i] * hi1
            c, s = _givens(H[:, j, j], np.abs(H[:, j + 1, j]))
            cs[:, j], sn[:, j] = c, s
            H[:, j, j] = c * H[:, j, j] + s * H[:, j + 1, j]
            H[:, j + 1, j] = 0.0
            g[:, j + 1] = -np.conj(s) * g[:, j]
            g[:, j] = c * g[:, j]
            res = np.abs(g[:, j + 1]) / np.where(bnorm > 0, bnorm, 1.0)
            steps[live] += 1
            iters[live] += 1
            for r in np.nonzero(live)[0]:
                history[r].append(float(res[r]))
            done = live & ((res <= tol) | (hn == 0) | (iters >= maxiter))
            live &= ~done
            if callback is not None:
                callback(j, res)
            if not live.any():
                break
        # x += V_k y with R_k y = g_k per column (its own step count)
        Y = np.zeros((R, m), dtype=hdt)
        for r in range(R):
            k = int(steps[r])
            if k == 0:
                continue
            Rk = H[r, :k, :k]
            Y[r, :k] = solve_triangular(Rk, g[r, :k], lower=False, check_finite=False)
        Yt = torch.as_tensor(Y.T.copy(), device=B.device).to(B.dtype)  # (m, R)
        X = X + torch.einsum("knr,kr->nr", Vs[:m], Yt)
        restarts += 1
    x = X[:, 0] if one else X
    return GMRESResult(x, iters, converged, history, restarts)


def fibonacci_sphere(n):
    """``n`` quasi-uniform points on the unit sphere (golden-angle spiral), equal weights ``4 pi / n``."""
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    th = math.pi * (1.0 + math.sqrt(5.0)) * i
    r = np.sqrt(np.maximum(1.0 - z * z, 0.0))
    X = np.stack([r * np.cos(th), r * np.sin(th), z], axis=1)
    return X, np.full(n, 4.0 * math.pi / n)


def laplace_sphere_dirichlet(n, *, source=(0.3, 2.0, -0.2), targets=None, tol=1e-8, leaf_size=200,
                             use_single_precision=False, gmres_tol=1e-10, restart=30, correction=True,
                             graph=True):
    """Interior Dirichlet problem on the unit sphere with data ``f = 1 / |x - source|`` (``|source| > 1``).

    Solves the double-layer equation with GMRES on the fast apply and evaluates the
    representation at interior ``targets`` (dense GPU blocks); returns a dict with the
    iteration count, residual history, the relative error against the exact harmonic
    function, and timings.
    """
    from . import operators as ops
    from .skel_fmm import SkelFMM

    X, w = fibonacci_sphere(n)
    x0 = np.asarray(source, dtype=np.float64)
    if np.linalg.norm(x0) <= 1.0:
        raise ValueError("the point source must lie outside the unit sphere")
    f = 1.0 / np.linalg.norm(X - x0, axis=1)
    if targets is None:
        targets = np.array([[0.1, 0.2, -0.3], [-0.4, 0.1, 0.2], [0.0, 0.0, 0.6], [0.5, -0.5, 0.1]])
    op = ops.LAPLACE_DOUBLE_LAYER_3D
    t0 = perf_counter()
    fmm = SkelFMM(X, op, geometry=X, tol=tol, leaf_size=leaf_size, use_single_precision=use_single_precision)
    t_build = perf_counter() - t0
    diag = 0.5 + (np.sqrt(w / math.pi) / 4.0 if correction else 0.0)
    A = FMMOperator(fmm, weights=w, diag=diag, graph=graph)
    dev = A.device
    b = torch.as_tensor(f, device=dev).to(A.dtype)
    torch.cuda.synchronize(dev)
    t0 = perf_counter()
    res = gmres(A, b, tol=gmres_tol, restart=restart, maxiter=10 * restart)
    torch.cuda.synchronize(dev)
    t_solve = perf_counter() - t0
    sigma = res.x.to(torch.float64)
    q = (torch.as_tensor(w, device=dev) * sigma)[None, :, None]
    xt = torch.as_tensor(targets, dtype=torch.float64, device=dev)[None]
    xs = torch.as_tensor(X, device=dev)[None]
    u = op.torch_apply(q, xt, xs, geom_src=xs).reshape(-1).cpu().numpy()
    exact = 1.0 / np.linalg.norm(targets - x0, axis=1)
    return {
        "n": int(n), "iterations": int(res.iterations[0]), "converged": bool(res.converged[0]),
        "residuals": res.residuals[0], "rel_err": float(np.linalg.norm(u - exact) / np.linalg.norm(exact)),
        "build_s": t_build, "solve_s": t_solve, "applies": A.napplies, "sigma": sigma.cpu().numpy(),
    }
