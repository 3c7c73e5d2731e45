"""Closed-form operator kernels and direct sums (oracle; test infrastructure only).

``kernel_matrix`` restates the numpy kernels of ``skelfmm/util.py:128-377``
(exact distances, zero where source and target coincide). ``kernel_matrix_torch``
restates the batched torch kernels of ``skelfmm/util_batched.py`` including their
distance clamps (``cdist`` 175-185: ``r^2 >= 1e-15`` in FP64, ``1e-8`` in FP32;
Laplace double layers 322-347 / 380-405: ``r^2 >= 1e-30`` / ``1e-12``) and the
self-pair diagonal mask (``set_self_interactions_to_value`` 154-172).
"""

from __future__ import annotations

import numpy as np
import torch

INV4PI = 1.0 / (4.0 * np.pi)

OPERATOR_NAMES = (
    "laplace_2d",
    "laplace_3d",
    "helmholtz_2d",
    "helmholtz_3d",
    "laplace_double_layer_3d",
    "laplace_adjoint_double_layer_3d",
    "helmholtz_double_layer_3d",
    "helmholtz_adjoint_double_layer_3d",
)
GEOM_SIDE = {
    "laplace_double_layer_3d": "source",
    "helmholtz_double_layer_3d": "source",
    "laplace_adjoint_double_layer_3d": "target",
    "helmholtz_adjoint_double_layer_3d": "target",
}
COMPLEX_OPS = {"helmholtz_2d", "helmholtz_3d", "helmholtz_double_layer_3d", "helmholtz_adjoint_double_layer_3d"}


def _value(name, r, rn, kappa):
    """K as a function of the distance r (> 0) and rn = (x - y) . n."""
    if name == "laplace_2d":
        return np.log(r)
    if name == "laplace_3d":
        return 1.0 / r
    if name == "helmholtz_2d":
        from scipy.special import j0, y0

        return j0(kappa * r) + 1j * y0(kappa * r)
    if name == "helmholtz_3d":
        return np.exp(1j * kappa * r) / r
    if name == "laplace_double_layer_3d":
        return -INV4PI * rn / r ** 3
    if name == "laplace_adjoint_double_layer_3d":
        return INV4PI * rn / r ** 3
    if name == "helmholtz_double_layer_3d":
        return INV4PI * rn * np.exp(1j * kappa * r) * (1j * kappa * r - 1.0) / r ** 3
    if name == "helmholtz_adjoint_double_layer_3d":
        return -INV4PI * rn * np.exp(1j * kappa * r) * (1j * kappa * r - 1.0) / r ** 3
    raise ValueError(name)


def kernel_matrix(name, X, Y, kappa=0.0, geom_trg=None, geom_src=None):
    """Dense K(X_i, Y_j) in FP64 with K = 0 where X_i == Y_j (``util.py:128-377``)."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    D = X[:, None, :] - Y[None, :, :]
    r = np.sqrt(np.sum(D * D, axis=-1))
    rn = np.zeros_like(r)
    side = GEOM_SIDE.get(name)
    if side == "source":
        rn = np.sum(D * np.asarray(geom_src)[None, :, :], axis=-1)
    elif side == "target":
        rn = np.sum(D * np.asarray(geom_trg)[:, None, :], axis=-1)
    zero = r == 0
    rs = np.where(zero, 1.0, r)
    K = _value(name, rs, rn, kappa)
    K = np.where(zero, 0.0, K)
    return K


def direct_sum(name, X, Y, q, kappa=0.0, geom_trg=None, geom_src=None, chunk=256):
    """``sum_j K(X_i, Y_j) q_j`` in FP64, chunked over targets."""
    q = np.asarray(q)
    out = None
    for a in range(0, X.shape[0], chunk):
        b = min(X.shape[0], a + chunk)
        gt = None if geom_trg is None else geom_trg[a:b]
        K = kernel_matrix(name, X[a:b], Y, kappa, geom_trg=gt, geom_src=geom_src)
        part = K @ q
        if out is None:
            out = np.zeros((X.shape[0],) + part.shape[1:], dtype=np.result_type(part.dtype, q.dtype))
        out[a:b] = part
    return out


def kernel_matrix_torch(name, XX, YY, kappa, self_pair, geom_trg=None, geom_src=None):
    """Batched dense blocks ``(B, m, n)`` with the torch clamps and self mask (``util_batched.py``)."""
    fp64 = XX.dtype == torch.float64
    diff = XX[..., :, None, :] - YY[..., None, :, :]
    r2 = torch.sum(diff * diff, dim=-1)
    if name in ("laplace_double_layer_3d", "laplace_adjoint_double_layer_3d"):
        r2 = torch.clamp(r2, min=1e-30 if fp64 else 1e-12)
    else:
        r2 = torch.clamp(r2, min=1e-15 if fp64 else 1e-8)
    r = torch.sqrt(r2)
    cdt = torch.complex128 if fp64 else torch.complex64
    side = GEOM_SIDE.get(name)
    rn = None
    if side == "source":
        rn = torch.sum(diff * geom_src[..., None, :, :], dim=-1)
    elif side == "target":
        rn = torch.sum(diff * geom_trg[..., :, None, :], dim=-1)
    if name == "laplace_2d":
        K = torch.log(r)
    elif name == "laplace_3d":
        K = 1.0 / r
    elif name == "helmholtz_2d":
        kr = kappa * r
        K = torch.complex(torch.special.bessel_j0(kr), torch.special.bessel_y0(kr))
    elif name == "helmholtz_3d":
        K = torch.exp(1j * kappa * r.to(cdt)) / r
    elif name == "laplace_double_layer_3d":
        K = -INV4PI * rn / (r2 * r)
    elif name == "laplace_adjoint_double_layer_3d":
        K = INV4PI * rn / (r2 * r)
    else:
        rc = r.to(cdt)
        f = torch.exp(1j * kappa * rc) * (1j * kappa * rc - 1.0) / (rc * rc * rc)
        K = INV4PI * rn.to(cdt) * f
        if name == "helmholtz_adjoint_double_layer_3d":
            K = -K
    if self_pair:
        m, n = K.shape[-2], K.shape[-1]
        eye = torch.eye(m, n, dtype=torch.bool)
        K = K.masked_fill(eye, 0.0)
    return K
