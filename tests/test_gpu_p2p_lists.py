"""Near-field kernels through the C ABI (`skf_p2p`) on hand-built lists.

Real builds give work items a few dozen source entries; the matrix-core near field keeps the
item's entry table in LDS up to 128 (LDS-direct staged kernel) or 256 (FP32 register-staged
kernel) entries and walks the global lists beyond that, and the symmetric kernel streams the
sources across entries up to 128 entries and stages entry by entry beyond. One group of 40
targets against NE single-point entries (every seventh subtracted) exercises both sides of
those limits, for both forms and for widths that take each staging path (64 and 100:
LDS-direct loads; 17 FP64 / 18 FP32: register staging), against a numpy FP64 sum of the same
Laplace 2D kernel (0.5 log r^2). The 3D Laplace cases (1/r) run the symmetric kernel in both
forms and at each inner-loop unroll on groups of up to 600 targets (several passes and work
items, partial lanes) against entries of several points; the 3D Laplace double layer and its
adjoint run it the same way (both directions' normal projections, FP64 rows in the precision
form) against the oracle's kernel matrices.
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


def _run(precision, width, X, centers, box, sign, q, symmetric=False, kernel=0, nt=NT, e_n=None, sym_unroll=0,
         normals=None, kappa=0.0, max_distance=0.0):
    NE = sign.shape[0]
    dim = X.shape[1]
    e_n = np.ones(NE, np.int64) if e_n is None else np.asarray(e_n, np.int64)
    e_pt = nt + np.concatenate(([0], np.cumsum(e_n)[:-1]))
    dev = torch.device("cuda")
    real = torch.float64 if precision == 0 else torch.float32
    P = X - centers[box, :dim] if precision == 1 else X  # the FP32 form stores box-centred points
    px = torch.as_tensor(P[:, 0], dtype=real, device=dev)
    py = torch.as_tensor(P[:, 1], dtype=real, device=dev)
    pz = torch.as_tensor(P[:, 2], dtype=real, device=dev) if dim == 3 else None
    geo = []
    if normals is not None:  # double layers: normals per point; the precision form's FP64 rows
        geo = [torch.as_tensor(normals[:, c], dtype=real, device=dev) for c in range(3)]
        if precision == 1:
            rows = np.concatenate([X.T, normals.T, np.sum(X * normals, axis=1)[None]])
            geo.append(torch.as_tensor(np.ascontiguousarray(rows), dtype=torch.float64, device=dev))
    qd = torch.as_tensor(q, dtype=real, device=dev).contiguous()
    ud = torch.zeros_like(qd)
    cen = torch.as_tensor(centers, dtype=torch.float64, device=dev)
    i32 = lambda a: torch.as_tensor(np.asarray(a, np.int32), device=dev)
    g = {"g_pt": i32([0]), "g_n": i32([nt]), "g_out": i32([0]), "g_box": i32([0])}
    e = {"e_pt": i32(e_pt), "e_n": i32(e_n), "e_q": i32(e_pt),
         "e_box": i32(1 + np.arange(NE)), "e_flags": i32(np.where(sign < 0, 1, 0))}
    # one work item per 512 targets (the symmetric kernel's item size; the one-sided kernels take
    # the group's first targets only in these tests)
    nitems = (nt + 511) // 512 if symmetric else 1
    items = i32([[0, 512 * i, 0, NE] for i in range(nitems)])
    a = hip.P2PArgs()
    keep = []
    if symmetric:
        # the whole list as pairs: targets get the sources' sums, every source the targets' sums
        one, sym = hip.P2PLists(), hip.P2PLists()
        one.ngroups = one.nentries = one.nitems = 0
        sym.ngroups, sym.nentries, sym.nitems = 1, NE, nitems
        for k, v in {**g, **e}.items():
            setattr(sym, k, hip.ptr(v))
        sym.items = hip.ptr(items)
        urev = torch.zeros(qd.numel(), dtype=torch.float64, device=dev)
        keep = [one, sym, urev]
        a.one, a.sym = ctypes.pointer(one), ctypes.pointer(sym)
        a.urev, a.urev_len = (hip.ptr(urev), urev.numel()) if precision == 1 else (None, 0)
    a.kernel, a.precision, a.width, a.kappa = kernel, precision, width, kappa
    a.max_distance = max_distance
    a.px, a.py, a.pz = hip.ptr(px), hip.ptr(py), (hip.ptr(pz) if pz is not None else None)
    a.sym_unroll = sym_unroll
    if normals is not None:
        a.nx, a.ny, a.nz = (hip.ptr(t) for t in geo[:3])
        if precision == 1:
            a.geo64, a.geo64_n = hip.ptr(geo[3]), int(X.shape[0])
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


def _case3(nt, ne, seed=0):
    """One group of nt targets (box 0) against ne entries of 1..40 points, each in a box of its
    own, every seventh entry subtracted; 3D points spread 0.02 around the box centres."""
    rng = np.random.default_rng(seed)
    centers = rng.random((ne + 1, 3))
    e_n = rng.integers(1, 41, size=ne)
    box = np.concatenate([np.zeros(nt, np.int64), np.repeat(1 + np.arange(ne), e_n)])
    X = centers[box] + 0.02 * rng.normal(size=(box.shape[0], 3))
    sign = np.where(np.arange(ne) % 7 == 3, -1.0, 1.0)
    return X, centers, box, sign, e_n


@pytest.mark.parametrize("precision,sym_unroll", [(0, 2), (1, 1), (1, 2), (1, 4)])
@pytest.mark.parametrize("nt,ne", [(40, 100), (600, 30), (97, 300), (1, 5), (20, 60), (70, 200), (280, 40)])
def test_symmetric_laplace3d_unrolls(precision, sym_unroll, nt, ne):
    """3D Laplace symmetric pairs: every inner-loop unroll of the FP32 form and the FP64 form
    against an FP64 sum, on one group of nt targets (one or two work items, passes of 256
    targets with partial lanes; last passes of 1, 20, 70, 88 and 24 targets in rows of 32 held by
    both wave halves) and ne entries of several points (up to 128 entries in the FP32 form:
    sources streamed across entries; beyond, and in the FP64 form: entry by entry)."""
    X, centers, box, sign, e_n = _case3(nt, ne)
    q = np.random.default_rng(3).normal(size=(X.shape[0], 1))
    d = X[:nt, None, :] - X[None, nt:, :]
    K = 1.0 / np.sqrt(np.sum(d * d, axis=-1)) * np.repeat(sign, e_n)[None, :]
    ref = np.zeros_like(q)
    ref[:nt] = K @ q[nt:]
    ref[nt:] = K.T @ q[:nt]
    print(f"3D symmetric precision {precision} unroll {sym_unroll}: {nt} targets, {ne} entries, "
          f"{X.shape[0] - nt} sources", flush=True)
    u = _run(precision, 1, X, centers, box, sign, q, symmetric=True, kernel=1, nt=nt, e_n=e_n, sym_unroll=sym_unroll)
    err_t = np.linalg.norm(u[:nt] - ref[:nt]) / np.linalg.norm(ref[:nt])
    err_s = np.linalg.norm(u[nt:] - ref[nt:]) / np.linalg.norm(ref[nt:])
    print(f"  rel-l2 targets {err_t:.2e}, sources {err_s:.2e}", flush=True)
    bar = 1e-13 if precision == 0 else 2e-6
    assert err_t < bar and err_s < bar


@pytest.mark.parametrize("kernel,name", [(4, "laplace_double_layer_3d"), (5, "laplace_adjoint_double_layer_3d")])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nt,ne", [(40, 100), (97, 300), (600, 30), (20, 60), (280, 150)])
def test_symmetric_double_layers(kernel, name, precision, nt, ne):
    """Double-layer pairs evaluated once for both directions: forward K(t, s), reverse K(s, t)
    (the normal on the other side), against the oracle's kernel matrices in FP64; 300 entries
    take the entry-by-entry staging, 100 the stream across entries (FP32 form)."""
    import torch

    from oracle.kernels import GEOM_SIDE, kernel_matrix_torch

    X, centers, box, sign, e_n = _case3(nt, ne, seed=5)
    nrm = np.random.default_rng(6).normal(size=X.shape)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    q = np.random.default_rng(7).normal(size=(X.shape[0], 1))
    sgn = np.repeat(sign, e_n)
    T, S = torch.as_tensor(X[:nt])[None], torch.as_tensor(X[nt:])[None]
    NT_, NS_ = torch.as_tensor(nrm[:nt])[None], torch.as_tensor(nrm[nt:])[None]
    if GEOM_SIDE[name] == "source":
        Kf = kernel_matrix_torch(name, T, S, 0.0, False, geom_src=NS_)[0].numpy()
        Kr = kernel_matrix_torch(name, S, T, 0.0, False, geom_src=NT_)[0].numpy()
    else:
        Kf = kernel_matrix_torch(name, T, S, 0.0, False, geom_trg=NT_)[0].numpy()
        Kr = kernel_matrix_torch(name, S, T, 0.0, False, geom_trg=NS_)[0].numpy()
    ref = np.zeros_like(q)
    ref[:nt] = (Kf * sgn[None, :]) @ q[nt:]
    ref[nt:] = (Kr * sgn[:, None]) @ q[:nt]
    print(f"{name} symmetric precision {precision}: {nt} targets, {ne} entries", flush=True)
    u = _run(precision, 1, X, centers, box, sign, q, symmetric=True, kernel=kernel, nt=nt, e_n=e_n, normals=nrm)
    err_t = np.linalg.norm(u[:nt] - ref[:nt]) / np.linalg.norm(ref[:nt])
    err_s = np.linalg.norm(u[nt:] - ref[nt:]) / np.linalg.norm(ref[nt:])
    print(f"  rel-l2 targets {err_t:.2e}, sources {err_s:.2e}", flush=True)
    bar = 1e-13 if precision == 0 else 2e-6
    assert err_t < bar and err_s < bar


@pytest.mark.parametrize("kappa", [40.0, 800.0])
def test_fp32_helmholtz_phase_over_the_hardware_range(kappa):
    """3D Helmholtz pairs in the FP32 form with the hardware sine/cosine of the unreduced phase in
    revolutions (t = kappa r / 2 pi up to about 220 at kappa 800, inside the hardware's |t| <= 256)
    against the same lists with sincosf (max_distance 0 turns the fast path off), both against an
    FP64 sum of e^{i kappa r} / r: the fast path loses nothing beyond the FP32 rounding of the
    phase that both paths share."""
    nt, ne = 200, 60
    X, centers, box, sign, e_n = _case3(nt, ne, seed=9)
    D = float(np.linalg.norm(X.max(axis=0) - X.min(axis=0)))
    assert kappa * D <= 1600
    rng = np.random.default_rng(10)
    qc = rng.normal(size=X.shape[0]) + 1j * rng.normal(size=X.shape[0])
    q = np.stack([qc.real, qc.imag], axis=1)
    d = X[:nt, None, :] - X[None, nt:, :]
    r = np.sqrt(np.sum(d * d, axis=-1))
    K = np.exp(1j * kappa * r) / r * np.repeat(sign, e_n)[None, :]
    ref = np.zeros(X.shape[0], np.complex128)
    ref[:nt] = K @ qc[nt:]
    ref[nt:] = K.T @ qc[:nt]
    out = {}
    for name, md in (("fast", D), ("sincosf", 0.0)):
        u = _run(1, 2, X, centers, box, sign, q, symmetric=True, kernel=3, nt=nt, e_n=e_n, kappa=kappa,
                 max_distance=md)
        out[name] = np.linalg.norm(u[:, 0] + 1j * u[:, 1] - ref) / np.linalg.norm(ref)
    print(f"kappa {kappa}: max phase {kappa * D / (2 * np.pi):.1f} revolutions, rel-l2 fast {out['fast']:.2e}, "
          f"sincosf {out['sincosf']:.2e}", flush=True)
    assert out["fast"] <= 1.5 * out["sincosf"] + 1e-7
    assert out["fast"] < 4 * 2.0**-24 * kappa * D + 1e-6  # a few FP32 roundings of the largest phase
