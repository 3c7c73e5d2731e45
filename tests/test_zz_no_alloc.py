"""GPU: ``skf_p2p_noalloc`` (near field with the library's own reverse-sum buffer) against
``skf_p2p`` (the caller's buffer), through the C ABI, on two calls in a row, the second on a
smaller input; and its argument checks.

The FP32 symmetric near field adds reverse sums with atomics, so two calls agree to FP32
rounding, not bit for bit: the bar is 1e-6 relative l2 on the potentials.
"""

import ctypes

import numpy as np
import pytest
import torch

from conftest import make_build, rel_l2
from skelfmm_amd import hip, operators as ops

pytestmark = pytest.mark.gpu

MAX_UREV = 1 << 22  # SKF_NOALLOC_MAX_UREV


def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    return torch.device("cuda", 0)


def _near_field_case(n, seed):
    """Charge and potential buffers of a Laplace 3D precision-form apply, and its near-field
    arguments with the symmetric split on."""
    from skelfmm_amd.batched import SkelFMMBatched

    X = np.random.default_rng(seed).random((n, 3))
    fmm = make_build(X, ops.LAPLACE_3D, tol=1e-6, leaf_size=150)
    dev = SkelFMMBatched(fmm.tree, ops.LAPLACE_3D).build_from_copy(fmm, use_single_precision=True)
    assert dev.use_symmetric
    x = torch.randn(n, dtype=torch.float32, device=_gpu())
    dev.apply_device(x)  # fills the plan's charge buffer (candidate and skeleton charges)
    plan = dev._plans[1]
    urev = torch.zeros(2 * dev.nslots, dtype=torch.float64, device=_gpu())
    args = dev._p2p_args(1, urev)
    assert args.sym and args.urev_len == 2 * dev.nslots
    return dev, plan, args, urev


def _run(entry, args, q, urev):
    u = torch.zeros_like(q)
    args.q = hip.ptr(q)
    args.u = hip.ptr(u)
    args.urev = None if urev is None else hip.ptr(urev)
    hip.check(entry(ctypes.byref(args), hip.stream_handle(q.device)))
    torch.cuda.synchronize()
    return u


def test_noalloc_matches_caller_buffer_on_two_calls():
    L = hip.lib()
    for n, seed in ((30000, 11), (8000, 12)):  # the second call on a smaller input
        print(f"skf_p2p_noalloc vs skf_p2p: Laplace 3D FP32 symmetric near field, N={n}", flush=True)
        dev, plan, args, urev = _near_field_case(n, seed)
        q = plan["qbuf"]
        ref = _run(L.skf_p2p, args, q, urev)
        got = _run(L.skf_p2p_noalloc, args, q, None)
        err = rel_l2(got.cpu().numpy(), ref.cpu().numpy())
        print(f"  rel-l2 {err:.2e}", flush=True)
        assert float(ref.abs().max()) > 0
        assert err <= 1e-6


def test_noalloc_rejects_bad_workspace_arguments():
    L = hip.lib()
    dev, plan, args, urev = _near_field_case(6000, 13)
    q = plan["qbuf"]
    u = torch.zeros_like(q)
    args.q, args.u = hip.ptr(q), hip.ptr(u)
    args.urev = hip.ptr(urev)  # a caller buffer is refused
    assert L.skf_p2p_noalloc(ctypes.byref(args), hip.stream_handle(q.device)) == 1
    assert b"no urev" in L.skf_last_error()
    args.urev = None
    args.urev_len = MAX_UREV + 1  # beyond the library's buffer
    assert L.skf_p2p_noalloc(ctypes.byref(args), hip.stream_handle(q.device)) == 1
    assert b"outside" in L.skf_last_error()
    args.urev_len = -1
    assert L.skf_p2p_noalloc(ctypes.byref(args), hip.stream_handle(q.device)) == 1
    torch.cuda.synchronize()
    assert float(u.abs().max()) == 0.0  # nothing ran
