"""The C-ABI library loads, exports every symbol the header declares, and rejects bad arguments (no GPU calls)."""

import ctypes
import os
import re

import pytest

from skelfmm_amd import hip

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "skelfmm_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(skf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_exported_symbols(hip_lib):
    names = _declared()
    assert set(names) == set(hip.EXPORTED_SYMBOLS)
    for n in names:
        assert hasattr(hip_lib, n), n


def test_version_and_error_string(hip_lib):
    assert hip_lib.skf_version() >= 1
    h = ctypes.c_void_p()
    rc = hip_lib.skf_plan_create(None, ctypes.byref(h))
    assert rc == 1
    assert b"null" in hip_lib.skf_last_error()
    assert hip_lib.skf_fill_zero(None, 0, 0, None) == 0
    assert hip_lib.skf_fill_zero(None, 10, 0, None) == 1


def test_plan_rejects_bad_descriptors(hip_lib):
    d = hip.PlanDesc()
    d.precision = 3
    d.width = 1
    h = ctypes.c_void_p()
    assert hip_lib.skf_plan_create(ctypes.byref(d), ctypes.byref(h)) == 1
    d.precision = 0
    d.width = 0
    assert hip_lib.skf_plan_create(ctypes.byref(d), ctypes.byref(h)) == 1
    d.width = 1
    d.qbuf = None
    assert hip_lib.skf_plan_create(ctypes.byref(d), ctypes.byref(h)) == 1
    a = hip.P2PArgs()
    a.nitems = 1
    a.kernel = 1
    a.width = 1
    assert hip_lib.skf_p2p(ctypes.byref(a), None) == 1  # null vectors


def test_blocks_rejects_mismatched_dimension(hip_lib):
    one = ctypes.c_void_p(1)
    # LAPLACE_3D (1) with ndim 2 is refused before any launch
    assert hip_lib.skf_p2p_blocks(1, 0, 1, 0.0, 1, 4, 4, 2, one, one, None, None, one, one, 0, None) == 1
    # source normals required for the double layer (4)
    assert hip_lib.skf_p2p_blocks(4, 0, 1, 0.0, 1, 4, 4, 3, one, one, None, None, one, one, 0, None) == 1


def test_apply_without_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present: the device path is exercised by the gpu tests")
    import numpy as np

    from skelfmm_amd import SkelFMM, operators

    with pytest.raises(RuntimeError, match="GPU"):
        SkelFMM(np.random.default_rng(0).random((200, 2)), operators.LAPLACE_2D, leaf_size=50)


def test_noalloc_near_field_rejects_workspace_arguments(hip_lib):
    """skf_p2p_noalloc refuses a caller buffer and reverse-sum lengths beyond its own buffer
    before any device work (no GPU needed)."""
    a = hip.P2PArgs()
    assert hip_lib.skf_p2p_noalloc(None, None) == 1
    fake = ctypes.c_void_p(16)  # never dereferenced: the checks come first
    a.q, a.u = fake, fake
    a.urev = fake
    assert hip_lib.skf_p2p_noalloc(ctypes.byref(a), None) == 1
    assert b"no urev" in hip_lib.skf_last_error()
    a.urev = None
    for bad in (-1, (1 << 22) + 1):
        a.urev_len = bad
        assert hip_lib.skf_p2p_noalloc(ctypes.byref(a), None) == 1
        assert b"outside" in hip_lib.skf_last_error()


def test_torch_operator_registers_apply(hip_lib):
    """The PyTorch operator library registers skelfmm::apply(int plan, Tensor x, Tensor(a!) y)
    for GPU tensors only, with the HIP library's source hash; CPU tensors are refused by the
    dispatcher before anything runs."""
    import torch

    from skelfmm_amd import _build

    ops = hip.torch_ops()
    assert _build.built_hash(hip.TORCH_LIB_PATH) == hip.build_id()
    schema = str(ops.apply.default._schema)
    assert schema.startswith("skelfmm::apply(int plan, Tensor x, Tensor(a!) y)")
    x = torch.zeros(4, dtype=torch.float64)
    with pytest.raises((NotImplementedError, RuntimeError)):
        ops.apply(1, x, torch.zeros_like(x))
