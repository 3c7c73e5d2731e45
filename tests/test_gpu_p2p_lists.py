"""Near-field kernels through the C ABI (`skf_p2p`) on hand-built lists.

Real builds give work items a few dozen source entries; the matrix-core near field keeps the
item's entry table in LDS up to 128 (LDS-direct staged kernel) or 256 (FP32 register-staged
kernel) entries and walks the global lists beyond that, and the symmetric kernel streams the
sources across entries up to 128 entries and stages entry by entry beyond. One group of 40
targets against NE single-point entries (every seventh subtracted) exercises both sides of
those limits, for both forms and for widths that take each staging path (64 and 100:
LDS-direct loads; 17 FP64 / 18 FP32: register staging), against a numpy FP64 sum of the same
Laplace 2D kernel (0.5 log r^2).
"""

import ctypes

import numpy as np
import pytest
import torch

from skelfmm_amd import hip

pytestmark = pytest.mark.gpu

NT, NE = 40, 300


def _case(seed=0, ne=NE):
    rng = np.random.default_rng(seed)
    NE = ne
    nbox = NE + 1
    centers = rng.random((nbox, 3))
    centers[:, 2] = 0.0
    box = np.concatenate([np.zeros(NT, np.int64), 1 + np.arange(NE)])
    X = centers[box, :2] + 0.01 * rng.normal(size=(NT + NE, 2))
    sign = np.where(np.arange(NE) % 7 == 3, -1.0, 1.0)
    return X, centers, box, sign


def _run(precision, width, X, centers, box, sign, q, symmetric=False):
    NE = sign.shape[0]
    dev = torch.device("cuda")
    real = torch.float64 if precision == 0 else torch.float32
    P = X - centers[box, :2] if precision == 1 else X  # the FP32 form stores box-centred points
    px = torch.as_tensor(P[:, 0], dtype=real, device=dev)
    py = torch.as_tensor(P[:, 1], dtype=real, device=dev)
    qd = torch.as_tensor(q, dtype=real, device=dev).contiguous()
    ud = torch.zeros_like(qd)
    cen = torch.as_tensor(centers, dtype=torch.float64, device=dev)
    i32 = lambda a: torch.as_tensor(np.asarray(a, np.int32), device=dev)
    g = {"g_pt": i32([0]), "g_n": i32([NT]), "g_out": i32([0]), "g_box": i32([0])}
    e = {"e_pt": i32(NT + np.arange(NE)), "e_n": i32(np.ones(NE)), "e_q": i32(NT + np.arange(NE)),
         "e_box": i32(1 + np.arange(NE)), "e_flags": i32(np.where(sign < 0, 1, 0))}
    items = i32([0, 0, 0, NE])
    a = hip.P2PArgs()
    keep = []
    if symmetric:
        # the whole list as pairs: targets get the sources' sums, every source the targets' sums
        one, sym = hip.P2PLists(), hip.P2PLists()
        one.ngroups = one.nentries = one.nitems = 0
        sym.ngroups, sym.nentries, sym.nitems = 1, NE, 1
        for k, v in {**g, **e}.items():
            setattr(sym, k, hip.ptr(v))
        sym.items = hip.ptr(items)
        urev = torch.zeros(qd.numel(), dtype=torch.float64, device=dev)
        keep = [one, sym, urev]
        a.one, a.sym = ctypes.pointer(one), ctypes.pointer(sym)
        a.urev, a.urev_len = (hip.ptr(urev), urev.numel()) if precision == 1 else (None, 0)
    a.kernel, a.precision, a.width, a.kappa = 0, precision, width, 0.0
    a.px, a.py, a.pz = hip.ptr(px), hip.ptr(py), None
    a.centers = hip.ptr(cen) if precision == 1 else None
    a.q, a.u = hip.ptr(qd), hip.ptr(ud)
    a.ngroups, a.nentries, a.nitems = 1, NE, 1
    for k, v in {**g, **e}.items():
        setattr(a, k, hip.ptr(v))
    a.items = hip.ptr(items)
    a.mfma_min_channels = 16
    hip.check(hip.lib().skf_p2p(ctypes.byref(a), hip.stream_handle(dev)))
    del keep
    torch.cuda.synchronize()
    return ud.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("precision,width", [(0, 64), (0, 100), (0, 17), (1, 64), (1, 100), (1, 18)])
def test_matrix_core_near_field_with_many_entries(precision, width):
    X, centers, box, sign = _case()
    rng = np.random.default_rng(1)
    q = rng.normal(size=(NT + NE, width))
    d = X[:NT, None, :] - X[None, NT:, :]
    K = 0.5 * np.log(np.maximum(np.sum(d * d, axis=-1), 1e-15)) * sign[None, :]
    ref = K @ q[NT:]
    u = _run(precision, width, X, centers, box, sign, q)
    err = np.linalg.norm(u[:NT] - ref) / np.linalg.norm(ref)
    print(f"precision {precision} width {width}: rel-l2 {err:.2e}", flush=True)
    assert err < (1e-13 if precision == 0 else 2e-6)
    assert np.all(u[NT:] == 0.0)  # only the group's targets receive sums


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("ne", [100, 300])
def test_symmetric_near_field_stream_and_entry_paths(precision, ne):
    """Symmetric pairs (one channel): 100 entries take the FP32 form's stream across entries,
    300 the entry-by-entry staging (the FP64 form always stages entry by entry)."""
    X, centers, box, sign = _case(ne=ne)
    q = np.random.default_rng(2).normal(size=(NT + ne, 1))
    d = X[:NT, None, :] - X[None, NT:, :]
    K = 0.5 * np.log(np.maximum(np.sum(d * d, axis=-1), 1e-15)) * sign[None, :]
    ref = np.zeros_like(q)
    ref[:NT] = K @ q[NT:]
    ref[NT:] = K.T @ q[:NT]
    u = _run(precision, 1, X, centers, box, sign, q, symmetric=True)
    err = np.linalg.norm(u - ref) / np.linalg.norm(ref)
    print(f"symmetric precision {precision} entries {ne}: rel-l2 {err:.2e}", flush=True)
    assert err < (1e-13 if precision == 0 else 2e-6)
