"""MI355X-native skeletonized FMM apply (drop-in for the ``skelfmm`` package).

``import skelfmm_amd as skelfmm`` gives the reference's public names
(``skelfmm/__init__.py:3-25``): ``SkelFMM``, ``RecursiveSkeletonization``,
``operators``, ``warmup_cuda_fmm_apply`` and the lazily imported
``SkelFMMBatched``.
"""

from . import operators
from .build import RecursiveSkeletonization
from .skel_fmm import SkelFMM, warmup_cuda_fmm_apply
from .tree import BalancedTree

__all__ = [
    "SkelFMM",
    "RecursiveSkeletonization",
    "SkelFMMBatched",
    "BalancedTree",
    "operators",
    "warmup_cuda_fmm_apply",
]


def __getattr__(name):
    if name == "SkelFMMBatched":
        from .batched import SkelFMMBatched

        return SkelFMMBatched
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
