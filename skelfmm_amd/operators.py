"""Kernel operator plugin objects (drop-in for ``skelfmm.operators``).

Mirrors ``skelfmm/operators.py:17-227``: the ``KernelOperator`` dataclass, the
``PointKernelOperator`` / ``GeometryAwareKernelOperator`` adapters, the eight
built-in operators and ``_require_operator``. Each built-in additionally carries
``kernel_id``, the operator code the HIP kernels dispatch on. ``torch_apply``
evaluates a batch of dense blocks on the GPU through the HIP block kernel
(``skf_p2p_blocks``), with the reference's signature (``util_batched.py``
``<kernel>(q, XX, YY, param, self_interaction, [src_geom|trg_geom])``).
``keops_func`` is always ``None``: there is no PyKeOps here.
"""

from dataclasses import dataclass
from functools import partial

from . import dense

__all__ = [
    "KernelOperator",
    "PointKernelOperator",
    "GeometryAwareKernelOperator",
    "LAPLACE_2D",
    "LAPLACE_3D",
    "HELMHOLTZ_2D",
    "HELMHOLTZ_3D",
    "LAPLACE_DOUBLE_LAYER_3D",
    "LAPLACE_ADJOINT_DOUBLE_LAYER_3D",
    "HELMHOLTZ_DOUBLE_LAYER_3D",
    "HELMHOLTZ_ADJOINT_DOUBLE_LAYER_3D",
    "BUILTIN_OPERATORS",
    "custom_operator",
]

# operator codes shared with skelfmm_amd/csrc/skelfmm_hip.hip (enum skf_kernel)
KID_LAPLACE_2D = 0
KID_LAPLACE_3D = 1
KID_HELMHOLTZ_2D = 2
KID_HELMHOLTZ_3D = 3
KID_LAPLACE_DL_3D = 4
KID_LAPLACE_ADL_3D = 5
KID_HELMHOLTZ_DL_3D = 6
KID_HELMHOLTZ_ADL_3D = 7


def _device_block_apply(kernel_id, q, x_trg, x_src, param, self_interaction, src_geom=None, trg_geom=None):
    from . import hip

    geom_trg = trg_geom
    geom_src = src_geom
    return hip.p2p_blocks(kernel_id, q, x_trg, x_src, param, self_interaction, geom_trg=geom_trg, geom_src=geom_src)


@dataclass
class KernelOperator:
    """Stateless kernel/operator interface for dense and batched evaluation."""

    name: str = "custom"
    ndim: int = -1
    numpy_func: object = None
    torch_func: object = None
    keops_func: object = None
    proxy_surface: str = "cube"
    geometry_side: str | None = None
    is_symmetric: bool = False
    kernel_id: int = -1
    is_complex: bool = False
    user_kernel: object = None  # jit.UserKernel of a custom_operator (run-time compiled near field)

    def __post_init__(self):
        if self.geometry_side not in {None, "source", "target"}:
            raise ValueError("geometry_side must be None, 'source', or 'target'")

    @property
    def requires_geometry(self):
        return self.geometry_side is not None

    @property
    def has_keops(self):
        return self.keops_func is not None

    def _geom_kwargs(self, geom_trg=None, geom_src=None):
        if self.geometry_side == "source":
            return {"src_geom": geom_src}
        if self.geometry_side == "target":
            return {"trg_geom": geom_trg}
        return {}

    def numpy_matrix(self, x_trg, x_src, *, geom_trg=None, geom_src=None, param=0.0,
                     assume_not_self_interaction=False):
        if self.numpy_func is None:
            raise NotImplementedError
        return self.numpy_func(
            x_trg,
            x_src,
            param=param,
            assume_not_self_interaction=assume_not_self_interaction,
            **self._geom_kwargs(geom_trg, geom_src),
        )

    def torch_apply(self, q, x_trg, x_src, *, geom_trg=None, geom_src=None, param=0.0,
                    self_interaction=False):
        if self.torch_func is None:
            raise NotImplementedError
        return self.torch_func(q, x_trg, x_src, param, self_interaction, **self._geom_kwargs(geom_trg, geom_src))

    def keops_apply(self, q, x_trg, x_src, *, geom_trg=None, geom_src=None, param=0.0,
                    self_interaction=False):
        if self.keops_func is None:
            return None
        geom_kwargs = self._geom_kwargs(geom_trg, geom_src)
        if param == 0.0:
            return self.keops_func(q, x_trg, x_src, self_interaction, **geom_kwargs)
        return self.keops_func(q, x_trg, x_src, param, self_interaction, **geom_kwargs)


class PointKernelOperator(KernelOperator):
    """Operator adapter for point kernels that do not use extra geometry."""

    def __init__(self, name, ndim, numpy_func, torch_func, keops_func=None, proxy_surface="cube",
                 is_symmetric=True, kernel_id=-1, is_complex=False):
        super().__init__(name=name, ndim=ndim, numpy_func=numpy_func, torch_func=torch_func,
                         keops_func=keops_func, proxy_surface=proxy_surface, is_symmetric=is_symmetric,
                         kernel_id=kernel_id, is_complex=is_complex)


class GeometryAwareKernelOperator(KernelOperator):
    """Operator for kernels that carry one aligned source or target geometry."""

    def __init__(self, name, ndim, numpy_func, torch_func, keops_func, *, geometry_side, kernel_id=-1,
                 is_complex=False):
        super().__init__(name=name, ndim=ndim, numpy_func=numpy_func, torch_func=torch_func,
                         keops_func=keops_func, proxy_surface="sphere", geometry_side=geometry_side,
                         kernel_id=kernel_id, is_complex=is_complex)


LAPLACE_2D = PointKernelOperator(
    name="laplace_2d", ndim=2, numpy_func=dense.laplace_2d,
    torch_func=partial(_device_block_apply, KID_LAPLACE_2D), kernel_id=KID_LAPLACE_2D)
LAPLACE_3D = PointKernelOperator(
    name="laplace_3d", ndim=3, numpy_func=dense.laplace_3d,
    torch_func=partial(_device_block_apply, KID_LAPLACE_3D), kernel_id=KID_LAPLACE_3D)
HELMHOLTZ_2D = PointKernelOperator(
    name="helmholtz_2d", ndim=2, numpy_func=dense.helmholtz_2d,
    torch_func=partial(_device_block_apply, KID_HELMHOLTZ_2D), kernel_id=KID_HELMHOLTZ_2D, is_complex=True)
HELMHOLTZ_3D = PointKernelOperator(
    name="helmholtz_3d", ndim=3, numpy_func=dense.helmholtz_3d,
    torch_func=partial(_device_block_apply, KID_HELMHOLTZ_3D), kernel_id=KID_HELMHOLTZ_3D, is_complex=True)
LAPLACE_DOUBLE_LAYER_3D = GeometryAwareKernelOperator(
    name="laplace_double_layer_3d", ndim=3, numpy_func=dense.laplace_3d_double_layer,
    torch_func=partial(_device_block_apply, KID_LAPLACE_DL_3D), keops_func=None,
    geometry_side="source", kernel_id=KID_LAPLACE_DL_3D)
LAPLACE_ADJOINT_DOUBLE_LAYER_3D = GeometryAwareKernelOperator(
    name="laplace_adjoint_double_layer_3d", ndim=3, numpy_func=dense.laplace_3d_adjoint_double_layer,
    torch_func=partial(_device_block_apply, KID_LAPLACE_ADL_3D), keops_func=None,
    geometry_side="target", kernel_id=KID_LAPLACE_ADL_3D)
HELMHOLTZ_DOUBLE_LAYER_3D = GeometryAwareKernelOperator(
    name="helmholtz_double_layer_3d", ndim=3, numpy_func=dense.helmholtz_3d_double_layer,
    torch_func=partial(_device_block_apply, KID_HELMHOLTZ_DL_3D), keops_func=None,
    geometry_side="source", kernel_id=KID_HELMHOLTZ_DL_3D, is_complex=True)
HELMHOLTZ_ADJOINT_DOUBLE_LAYER_3D = GeometryAwareKernelOperator(
    name="helmholtz_adjoint_double_layer_3d", ndim=3, numpy_func=dense.helmholtz_3d_adjoint_double_layer,
    torch_func=partial(_device_block_apply, KID_HELMHOLTZ_ADL_3D), keops_func=None,
    geometry_side="target", kernel_id=KID_HELMHOLTZ_ADL_3D, is_complex=True)

BUILTIN_OPERATORS = (
    LAPLACE_2D,
    LAPLACE_3D,
    HELMHOLTZ_2D,
    HELMHOLTZ_3D,
    LAPLACE_DOUBLE_LAYER_3D,
    LAPLACE_ADJOINT_DOUBLE_LAYER_3D,
    HELMHOLTZ_DOUBLE_LAYER_3D,
    HELMHOLTZ_ADJOINT_DOUBLE_LAYER_3D,
)


def _user_numpy(uk, XX, YY, param=0.0, assume_not_self_interaction=False):
    """Dense ``K(x_i, y_j)`` of a user formula; zero where the points coincide."""
    import numpy as np

    XX = np.asarray(XX, dtype=np.float64)
    YY = np.asarray(YY, dtype=np.float64)
    if XX.shape[-1] != uk.ndim or YY.shape[-1] != uk.ndim:
        raise ValueError(f"points must have {uk.ndim} coordinates")
    diff, r2 = dense._diff_r2(XX, YY)
    if assume_not_self_interaction:
        return np.array(uk.numpy_values(diff, r2, param))
    zero = r2 == 0
    v = np.array(uk.numpy_values(diff, np.where(zero, 1.0, r2), param))
    v[zero] = 0.0
    return v


def _user_torch(uk, q, XX, YY, param, self_interaction):
    """``(B, m, G) = K(XX, YY) @ q`` for a user formula in torch (any device)."""
    import torch

    squeeze = XX.dim() == 2
    if squeeze:
        XX, YY, q = XX[None], YY[None], q[None]
    diff = XX[..., :, None, :] - YY[..., None, :, :]
    r2 = torch.sum(diff * diff, dim=-1)
    zero = r2 == 0
    K = uk.torch_values(diff, torch.where(zero, torch.ones_like(r2), r2), param)
    K = torch.where(zero, torch.zeros_like(K), K)
    if self_interaction:
        m, n = K.shape[-2], K.shape[-1]
        K = K.masked_fill(torch.eye(m, n, dtype=torch.bool, device=K.device), 0.0)
    out = torch.matmul(K, q.to(K.dtype))
    return out[0] if squeeze else out


def custom_operator(name, expression, ndim=3, symmetric=None):
    """A user-defined real point kernel ``K(x, y) = expression`` (``skelfmm_amd.jit`` language).

    The host build evaluates it with numpy, ``torch_apply`` with torch, and the device apply
    compiles its near field at run time (hiprtc) into the library's one-sided near-field
    kernel. ``symmetric`` (default: the formula uses no ``dx dy dz``, so ``K(x, y) = K(y, x)``)
    selects the single proxy block of the build, as for the built-in single layers."""
    from . import jit

    uk = jit.UserKernel(expression, ndim)
    if symmetric is None:
        symmetric = not ({"dx", "dy", "dz"} & uk.names_used)
    op = PointKernelOperator(name=str(name), ndim=int(ndim), numpy_func=partial(_user_numpy, uk),
                             torch_func=partial(_user_torch, uk), is_symmetric=bool(symmetric),
                             kernel_id=uk.kernel_id)
    op.user_kernel = uk
    return op


def _require_operator(operator):
    """Validate that an explicit `KernelOperator` was provided."""
    if not isinstance(operator, KernelOperator):
        raise TypeError(
            "Expected a KernelOperator instance. "
            "Use objects from skelfmm_amd.operators such as operators.LAPLACE_3D."
        )
    return operator
