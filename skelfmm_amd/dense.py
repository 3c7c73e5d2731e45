"""Dense host kernel matrices for the eight built-in operators (numpy, FP64).

These are the operators' ``numpy_func`` entries. The host build uses them for
the proxy matrices, and ``SkelFMM.relerr_fmm_apply_check`` uses them for the
subset direct sum. They follow ``skelfmm/util.py:128-377``: no ``1/(2*pi)`` or
``1/(4*pi)`` factor on the single-layer kernels, a ``1/(4*pi)`` factor and the
signs of the code (not of the docstrings) on the double-layer kernels, and a
zero value where source and target coincide.
"""

from __future__ import annotations

import numpy as np
from scipy.special import j0, y0

__all__ = [
    "laplace_2d",
    "laplace_3d",
    "helmholtz_2d",
    "helmholtz_3d",
    "laplace_3d_double_layer",
    "laplace_3d_adjoint_double_layer",
    "helmholtz_3d_double_layer",
    "helmholtz_3d_adjoint_double_layer",
]

_INV4PI = 1.0 / (4.0 * np.pi)


def _diff_r2(XX, YY):
    XX = np.asarray(XX, dtype=np.float64)
    YY = np.asarray(YY, dtype=np.float64)
    diff = XX[:, None, :] - YY[None, :, :]
    r2 = np.einsum("ijk,ijk->ij", diff, diff)
    return diff, r2


def laplace_2d(XX, YY, param=0, assume_not_self_interaction=False):
    """``log|x-y|`` (``util.py:128-149``); zero at coincident points."""
    assert param == 0, "Parameter must be 0 for Laplace kernel"
    assert XX.shape[-1] == 2 and YY.shape[-1] == 2
    _, r2 = _diff_r2(XX, YY)
    if not assume_not_self_interaction:
        r2[r2 == 0] = 1.0
    return 0.5 * np.log(r2)


def laplace_3d(XX, YY, param=0, assume_not_self_interaction=False):
    """``1/|x-y|`` (``util.py:151-174``); zero at coincident points."""
    assert param == 0, "Parameter must be 0 for Laplace kernel"
    assert XX.shape[-1] == 3 and YY.shape[-1] == 3
    _, r2 = _diff_r2(XX, YY)
    r = np.sqrt(r2)
    if assume_not_self_interaction:
        return 1.0 / r
    out = np.zeros_like(r)
    nz = r != 0
    out[nz] = 1.0 / r[nz]
    return out


def helmholtz_2d(XX, YY, param, assume_not_self_interaction=False):
    """``J0(k r) + i Y0(k r)`` (``util.py:176-202``); zero at coincident points."""
    _, r2 = _diff_r2(XX, YY)
    r = np.sqrt(r2)
    zero = None
    if not assume_not_self_interaction:
        zero = r == 0
        r[zero] = 1.0
    kr = param * r
    out = j0(kr) + 1j * y0(kr)
    if zero is not None:
        out[zero] = 0
    return out


def helmholtz_3d(XX, YY, param, assume_not_self_interaction=False):
    """``exp(i k r) / r`` (``util.py:204-232``); zero at coincident points."""
    _, r2 = _diff_r2(XX, YY)
    r = np.sqrt(r2)
    if assume_not_self_interaction:
        return np.exp(1j * param * r) / r
    out = np.zeros(r.shape, dtype=np.complex128)
    nz = r != 0
    out[nz] = np.exp(1j * param * r[nz]) / r[nz]
    return out


def _dl_parts(XX, YY, normals, side, assume_not_self_interaction):
    diff, r2 = _diff_r2(XX, YY)
    if not assume_not_self_interaction:
        r2 = r2.copy()
        r2[r2 == 0] = 1.0
    if side == "source":
        rn = np.einsum("ijk,jk->ij", diff, np.asarray(normals, dtype=np.float64))
    else:
        rn = np.einsum("ijk,ik->ij", diff, np.asarray(normals, dtype=np.float64))
    r = np.sqrt(r2)
    return rn, r, r2


def laplace_3d_double_layer(XX, YY, param=0, src_geom=None, assume_not_self_interaction=False):
    """``-(1/4pi) ((x-y).n_y) / r^3`` (``util.py:234-265``)."""
    assert param == 0, "Parameter must be 0 for Laplace double-layer kernel"
    if src_geom is None:
        raise ValueError("src_geom is required for the Laplace double-layer kernel")
    rn, r, r2 = _dl_parts(XX, YY, src_geom, "source", assume_not_self_interaction)
    return -_INV4PI * rn / (r2 * r)


def laplace_3d_adjoint_double_layer(XX, YY, param=0, trg_geom=None, assume_not_self_interaction=False):
    """``+(1/4pi) ((x-y).n_x) / r^3`` (``util.py:267-298``)."""
    assert param == 0, "Parameter must be 0 for Laplace adjoint double-layer kernel"
    if trg_geom is None:
        raise ValueError("trg_geom is required for the Laplace adjoint double-layer kernel")
    rn, r, r2 = _dl_parts(XX, YY, trg_geom, "target", assume_not_self_interaction)
    return _INV4PI * rn / (r2 * r)


def helmholtz_3d_double_layer(XX, YY, param, src_geom=None, assume_not_self_interaction=False):
    """``(1/4pi) ((x-y).n_y) e^{ikr} (ikr-1) / r^3`` (``util.py:301-338``)."""
    assert param > 0, "Parameter must be positive for Helmholtz double-layer kernel"
    if src_geom is None:
        raise ValueError("src_geom is required for the Helmholtz double-layer kernel")
    rn, r, r2 = _dl_parts(XX, YY, src_geom, "source", assume_not_self_interaction)
    kr = param * r
    return _INV4PI * rn * (np.exp(1j * kr) * (1j * kr - 1.0) / (r2 * r))


def helmholtz_3d_adjoint_double_layer(XX, YY, param, trg_geom=None, assume_not_self_interaction=False):
    """``-(1/4pi) ((x-y).n_x) e^{ikr} (ikr-1) / r^3`` (``util.py:341-377``)."""
    assert param > 0, "Parameter must be positive for Helmholtz adjoint double-layer kernel"
    if trg_geom is None:
        raise ValueError("trg_geom is required for the Helmholtz adjoint double-layer kernel")
    rn, r, r2 = _dl_parts(XX, YY, trg_geom, "target", assume_not_self_interaction)
    kr = param * r
    return -_INV4PI * rn * (np.exp(1j * kr) * (1j * kr - 1.0) / (r2 * r))
