"""Host-side pieces of the ``SkelFMM`` wrapper (no GPU): dtype rules, storage and rank
accounting on a host build, the deferred-build state and argument checks."""

import numpy as np
import pytest
import torch

from conftest import make_build
from skelfmm_amd import SkelFMM, operators as ops, warmup_cuda_fmm_apply
from skelfmm_amd.skel_fmm import _batched_tensor_dtype, _host_build_bytes, _largest_rank


def test_vector_dtype_rule():
    """Reference skel_fmm.py:32-35: real input keeps the stored real type, complex input takes
    the complex type of the stored width."""
    r, c = np.zeros(3), np.zeros(3, complex)
    assert _batched_tensor_dtype(r, torch.float64) == torch.float64
    assert _batched_tensor_dtype(r, torch.float32) == torch.float32
    assert _batched_tensor_dtype(c, torch.float64) == torch.complex128
    assert _batched_tensor_dtype(c, torch.float32) == torch.complex64
    assert _batched_tensor_dtype(r.astype(np.float32), torch.float64) == torch.float64


def test_storage_and_rank_accounting_on_a_host_build():
    X = np.random.default_rng(0).random((3000, 2))
    f = make_build(X, ops.LAPLACE_2D, tol=1e-6, leaf_size=64)
    expect = f.nbytes_proj + sum(c.nbytes for c in f.cand)
    expect += f.bs_list.nbytes + f.rank_list.nbytes + f.tree_boxes.nbytes + f.tree_lev_sep.nbytes
    assert _host_build_bytes(f) == expect > 0
    assert _host_build_bytes(None) == 0
    assert _largest_rank(f) == int(f.rank_list.max()) > 0
    assert _largest_rank(None) == 0


def test_deferred_build_state_and_arguments():
    X = np.random.default_rng(1).random((500, 3))
    fmm = SkelFMM(X, ops.LAPLACE_3D, tol=1e-6, rank_or_tol=1e-7, leaf_size=50, defer_build=True)
    assert fmm.rank_or_tol == 1e-7 and fmm.leaf_size == 50
    assert fmm.build_time_seconds == 0.0 and fmm.nbytes == 0 and fmm.kernel_backend is None
    assert not fmm._is_built
    with pytest.raises(ValueError):
        SkelFMM(X, ops.LAPLACE_3D, defer_build=True, build_device="tpu")
    with pytest.raises(TypeError):
        # the extensions are keyword-only: a 16th positional argument is refused
        SkelFMM(X, ops.LAPLACE_3D, None, 1e-5, None, 50, False, 0.0, True, False, True, 500, "auto", -1, -1, False)


def test_warmup_skips_host_operators():
    class HostOnly:
        device = torch.device("cpu")

        def apply(self, x):
            raise AssertionError("must not be called")

    assert warmup_cuda_fmm_apply(HostOnly(), np.zeros(3)) is False
