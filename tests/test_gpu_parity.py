"""GPU parity: the HIP apply against the CPU oracle on identical build data.

Bars: FP64 apply within 1e-12 relative l2 of the oracle's padded restatement of
``SkelFMMBatched.matvec`` (``skelfmm/batched_fmm.py:539-564``) on the same
build; FP32 form within the error stated per test against the FP64 direct sum.
"""

import numpy as np
import pytest
import torch

from conftest import annulus_points, make_build, rel_l2, sphere_points
from oracle.batched_apply import BatchedApplyCPU
from oracle.direct import direct_sum_c
from oracle.kernels import GEOM_SIDE, kernel_matrix_torch
from skelfmm_amd import operators as ops

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    return torch.device("cuda", 0)


def _device_apply(fmm, q, single=False):
    from skelfmm_amd.batched import SkelFMMBatched

    dev = SkelFMMBatched(fmm.tree, fmm.operator, kappa=fmm.kappa).build_from_copy(fmm, use_single_precision=single)
    y = dev.matvec(torch.as_tensor(q).to(_gpu()))
    torch.cuda.synchronize()
    return y.cpu().numpy(), dev


def _case(name):
    rng = np.random.default_rng(0)
    if name == "laplace_2d_square":
        return rng.random((5000, 2)), ops.LAPLACE_2D, 0.0, None, 64
    if name == "laplace_2d_annulus":
        return annulus_points(6000), ops.LAPLACE_2D, 0.0, None, 64
    if name == "laplace_3d_cube":
        return rng.random((12000, 3)), ops.LAPLACE_3D, 0.0, None, 150
    if name == "helmholtz_3d_cube":
        return rng.random((4000, 3)), ops.HELMHOLTZ_3D, 5.0, None, 100
    if name == "helmholtz_2d_square":
        return rng.random((3000, 2)), ops.HELMHOLTZ_2D, 6.0, None, 64
    v = sphere_points(5000)
    if name == "laplace_dl_sphere":
        return v, ops.LAPLACE_DOUBLE_LAYER_3D, 0.0, v, 100
    if name == "laplace_adl_sphere":
        return v, ops.LAPLACE_ADJOINT_DOUBLE_LAYER_3D, 0.0, v, 100
    if name == "helmholtz_dl_sphere":
        return v, ops.HELMHOLTZ_DOUBLE_LAYER_3D, 4.0, v, 100
    if name == "helmholtz_adl_sphere":
        return v, ops.HELMHOLTZ_ADJOINT_DOUBLE_LAYER_3D, 4.0, v, 100
    raise KeyError(name)


CASES = ["laplace_2d_square", "laplace_2d_annulus", "laplace_3d_cube", "helmholtz_3d_cube", "helmholtz_2d_square",
         "laplace_dl_sphere", "laplace_adl_sphere", "helmholtz_dl_sphere", "helmholtz_adl_sphere"]


@pytest.mark.parametrize("name", CASES)
def test_fp64_apply_matches_oracle(name):
    X, op, kappa, geom, leaf = _case(name)
    print(f"case {name} N={X.shape[0]}", flush=True)
    fmm = make_build(X, op, kappa=kappa, tol=1e-6, leaf_size=leaf, geometry=geom)
    q = np.random.RandomState(1).randn(X.shape[0])
    y, _ = _device_apply(fmm, q)
    ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(q).numpy()
    if op is ops.HELMHOLTZ_2D:
        # torch's bessel_y0 (the oracle's batched kernels) and the device's y0 differ near 1e-8;
        # compare with the serial restatement that uses scipy's Bessel functions
        from oracle.serial_apply import SerialApply

        ref = SerialApply(fmm).matvec(q)
    err = rel_l2(y, ref)
    print(f"  rel-l2 vs oracle = {err:.3e}", flush=True)
    assert err <= FP64_TOL


@pytest.mark.parametrize("kid,name", [(o.kernel_id, o.name) for o in ops.BUILTIN_OPERATORS])
@pytest.mark.parametrize("self_pair", [False, True])
def test_p2p_blocks_matches_torch_kernels(kid, name, self_pair):
    from skelfmm_amd import hip

    _gpu()
    rng = np.random.default_rng(kid)
    d = 2 if name.endswith("2d") else 3
    B, m = 3, 37
    n = m if self_pair else 53
    xt = rng.random((B, m, d))
    xs = xt.copy() if self_pair else rng.random((B, n, d)) + 0.3
    g = rng.normal(size=(B, max(m, n), 3))
    g /= np.linalg.norm(g, axis=-1, keepdims=True)
    gt = g[:, :m] if GEOM_SIDE.get(name) == "target" else None
    gs = g[:, :n] if GEOM_SIDE.get(name) == "source" else None
    kappa = 3.0 if "helmholtz" in name else 0.0
    q = rng.normal(size=(B, n, 2))
    if "helmholtz" in name:
        q = q + 1j * rng.normal(size=(B, n, 2))
    tt = lambda a: None if a is None else torch.as_tensor(a)
    K = kernel_matrix_torch(name, tt(xt), tt(xs), kappa, self_pair, geom_trg=tt(gt), geom_src=tt(gs))
    ref = (K.to(torch.as_tensor(q).dtype) if not K.is_complex() else K) @ torch.as_tensor(q)
    dev = _gpu()
    out = hip.p2p_blocks(kid, tt(q).to(dev), tt(xt).to(dev), tt(xs).to(dev), kappa, self_pair,
                         geom_trg=None if gt is None else tt(gt).to(dev),
                         geom_src=None if gs is None else tt(gs).to(dev))
    err = rel_l2(out.cpu().numpy(), ref.numpy())
    tol = 1e-7 if name == "helmholtz_2d" else 1e-13
    assert err <= tol, err


def test_precision_form_against_direct_sum():
    rng = np.random.default_rng(0)
    X = rng.random((30000, 3))
    fmm = make_build(X, ops.LAPLACE_3D, tol=1e-6, leaf_size=200)
    q = np.random.RandomState(1).randn(X.shape[0])
    idx = np.random.default_rng(0).choice(X.shape[0], 1000, replace=False)
    ref = direct_sum_c("laplace_3d", X[idx], X, q)
    y64, _ = _device_apply(fmm, q, single=False)
    y32, dev32 = _device_apply(fmm, q.astype(np.float32), single=True)
    assert y32.dtype == np.float32
    e64 = rel_l2(y64[idx], ref)
    e32 = rel_l2(y32[idx].astype(np.float64), ref)
    print(f"laplace_3d N=30000 tol=1e-6: FP64 err {e64:.3e}  FP32 err {e32:.3e}", flush=True)
    assert e64 < 1e-5
    assert e32 < 2e-6
    assert rel_l2(y32, y64) < 2e-6


def test_fp32_helmholtz_fast_trig_accuracy():
    """The FP32 form's Helmholtz near field with the hardware sine/cosine of the phase in
    revolutions (taken when kappa times the cloud diameter <= 1600) against the same build with
    the library's sincosf (max_distance 0 turns the fast path off) and the FP64 apply, at kappa
    40 (kernel arguments up to ~50 in the near field)."""
    from skelfmm_amd.batched import SkelFMMBatched

    X = np.random.default_rng(5).random((8000, 3))
    fmm = make_build(X, ops.HELMHOLTZ_3D, kappa=40.0, tol=1e-6, leaf_size=100)
    rs = np.random.RandomState(4)
    q = rs.randn(X.shape[0]) + 1j * rs.randn(X.shape[0])
    x = torch.as_tensor(q).to(_gpu())
    y64 = SkelFMMBatched(fmm.tree, fmm.operator, kappa=fmm.kappa).build_from_copy(fmm).matvec(x).cpu().numpy()
    d32 = SkelFMMBatched(fmm.tree, fmm.operator, kappa=fmm.kappa).build_from_copy(fmm, use_single_precision=True)
    assert 0 < 40.0 * d32.max_distance <= 1600
    x32 = torch.as_tensor(q.astype(np.complex64)).to(_gpu())
    y_fast = d32.matvec(x32).cpu().numpy().astype(np.complex128)
    d32.max_distance = 0.0
    for k in list(d32._plans):
        d32._destroy_plan(k)
    d32._graphs = {}
    y_lib = d32.matvec(x32).cpu().numpy().astype(np.complex128)
    e_fast, e_lib = rel_l2(y_fast, y64), rel_l2(y_lib, y64)
    print(f"helmholtz kappa 40 FP32: fast trig {e_fast:.2e}, sincosf {e_lib:.2e}", flush=True)
    assert e_fast <= 1.5 * e_lib + 1e-7
    assert e_fast < 2e-6


def test_precision_form_helmholtz_complex64():
    """Mirrors test_batched.py:469-495: complex64 result within 5e-4 of FP64."""
    rng = np.random.default_rng(3)
    X = rng.random((3000, 3))
    fmm = make_build(X, ops.HELMHOLTZ_3D, kappa=2.0, tol=1e-5, leaf_size=80)
    q = np.random.default_rng(0).random(X.shape[0]).astype(np.float32)
    y32, dev = _device_apply(fmm, q, single=True)
    assert y32.dtype == np.complex64
    ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(q.astype(np.float64)).numpy()
    assert rel_l2(y32.astype(np.complex128), ref) < 5e-4


@pytest.mark.parametrize("single", [False, True])
def test_multi_rhs_equals_columns(single):
    rng = np.random.default_rng(5)
    X = rng.random((4000, 2))
    fmm = make_build(X, ops.LAPLACE_2D, tol=1e-6, leaf_size=64)
    Q = np.random.RandomState(2).randn(X.shape[0], 5)
    if single:
        Q = Q.astype(np.float32)
    Y, _ = _device_apply(fmm, Q, single=single)
    cols = np.stack([_device_apply(fmm, Q[:, c].copy(), single=single)[0] for c in range(Q.shape[1])], axis=1)
    assert rel_l2(Y, cols) < (1e-6 if single else 1e-13)
    if not single:
        ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(Q).numpy()
        assert rel_l2(Y, ref) <= FP64_TOL


@pytest.mark.parametrize("nrhs", [3, 16])
def test_multi_rhs_complex(nrhs):
    X = np.random.default_rng(6).random((3000, 3))
    fmm = make_build(X, ops.HELMHOLTZ_3D, kappa=3.0, tol=1e-6, leaf_size=100)
    rs = np.random.RandomState(3)
    Q = rs.randn(X.shape[0], nrhs) + 1j * rs.randn(X.shape[0], nrhs)
    Y, _ = _device_apply(fmm, Q)
    ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(Q).numpy()
    assert rel_l2(Y, ref) <= FP64_TOL


def test_config1_full_size_parity():
    """Plain 2D Laplace, N=100,000 uniform (default_rng(0)), RandomState(1) charges, tol 1e-5, leaf 200."""
    X = np.random.default_rng(0).random((100000, 2))
    q = np.random.RandomState(1).randn(X.shape[0])
    fmm = make_build(X, ops.LAPLACE_2D, tol=1e-5, leaf_size=200)
    y, _ = _device_apply(fmm, q)
    ref = BatchedApplyCPU(fmm).matvec(q).numpy()
    err = rel_l2(y, ref)
    print(f"config1 N=100000: GPU vs oracle rel-l2 = {err:.3e}", flush=True)
    assert err <= FP64_TOL
    idx = np.random.default_rng(0).choice(X.shape[0], 1000, replace=False)
    dref = direct_sum_c("laplace_2d", X[idx], X, q)
    assert rel_l2(y[idx], dref) < 1e-5


@pytest.mark.parametrize("n", [1, 2, 7, 64, 65])
def test_small_and_single_leaf_inputs(n):
    X = np.random.default_rng(n).random((n, 3))
    q = np.random.RandomState(n).randn(n)
    fmm = make_build(X, ops.LAPLACE_3D, tol=1e-6, leaf_size=64)
    y, _ = _device_apply(fmm, q)
    ref = direct_sum_c("laplace_3d", X, X, q)
    assert rel_l2(y, ref) <= 1e-13 or np.allclose(y, ref, atol=1e-14)


def test_coincident_points_follow_batched_clamp():
    """Two distinct points at the same location: the batched kernels clamp r^2 (util_batched.py:175-185)."""
    X = np.random.default_rng(9).random((3000, 2))
    X[17] = X[1234]
    q = np.random.RandomState(9).randn(X.shape[0])
    fmm = make_build(X, ops.LAPLACE_2D, tol=1e-6, leaf_size=64)
    y, _ = _device_apply(fmm, q)
    ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(q).numpy()
    assert rel_l2(y, ref) <= FP64_TOL


def test_skelfmm_public_api(capsys):
    from skelfmm_amd import SkelFMM

    X = np.random.default_rng(0).random((3000, 2))
    fmm = SkelFMM(X, ops.LAPLACE_2D, tol=1e-6, leaf_size=64, defer_build=True)
    assert fmm.build_time_seconds == 0.0
    x = np.random.RandomState(0).randn(X.shape[0])
    y = fmm.apply(x)
    assert "Building SkelFMM on first matvec" in capsys.readouterr().out
    assert y.shape == (X.shape[0],) and y.dtype == np.float64
    assert fmm.kernel_backend == "hip" and fmm.device.type == "cuda"
    assert fmm.rank_max > 0 and fmm.nbytes > 0
    assert fmm.relerr_fmm_apply_check(x) < 1e-6
    b = fmm.benchmark_matvec(nrepeat=2, warmup=1)
    assert b["min_seconds"] > 0
    with pytest.raises(ValueError):
        fmm.apply(np.zeros((X.shape[0], 2)))
    single = SkelFMM(X, ops.LAPLACE_2D, tol=1e-6, leaf_size=64, use_single_precision=True)
    assert single.apply(x).dtype == np.float32
    hz = SkelFMM(X[:1500], ops.HELMHOLTZ_2D, tol=1e-6, leaf_size=64, param=4.0)
    assert hz.apply(x[:1500]).dtype == np.complex128


def _annulus_instances(ns, seed=2):
    from conftest import annulus_points

    return [annulus_points(n, seed + i) for i, n in enumerate(ns)]


def test_batched_instances_match_separate_applies():
    from skelfmm_amd.batched import BatchedInstances

    pts = _annulus_instances((3000, 9000, 1200, 5000))
    builds = [make_build(P, ops.LAPLACE_2D, tol=1e-5, leaf_size=64) for P in pts]
    rng = np.random.default_rng(0)
    qs = [rng.normal(size=(P.shape[0], 3)) for P in pts]
    dev = BatchedInstances(builds)
    y = dev.matvec(dev.concat(qs))
    torch.cuda.synchronize()
    for part, b, q in zip(dev.split(y), builds, qs):
        ref = BatchedApplyCPU(b, bytes_budget=2e8).matvec(q).numpy()
        assert rel_l2(part.cpu().numpy(), ref) <= FP64_TOL


def test_batched_instances_of_gpu_builds():
    """Instances built on the GPU (tree, batched IDs, device work tables) and merged: each
    instance's part of the one apply matches the oracle on that instance's build data."""
    from skelfmm_amd.batched import BatchedInstances
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.gpu_tree import GPUBalancedTree

    pts = _annulus_instances((3000, 9000, 1200, 5000))
    builds = []
    for P in pts:
        f = RecursiveSkeletonization(GPUBalancedTree(P, leaf_size=64), ops.LAPLACE_2D, kappa=0.0, tol=1e-5)
        f.skel_tree(npoints_max=500, device="gpu")
        f.setup_lists()
        builds.append(f)
    rng = np.random.default_rng(0)
    qs = [rng.normal(size=(P.shape[0], 3)) for P in pts]
    dev = BatchedInstances(builds)
    y = dev.matvec(dev.concat(qs))
    torch.cuda.synchronize()
    for part, b, q in zip(dev.split(y), builds, qs):
        ref = BatchedApplyCPU(b, bytes_budget=2e8).matvec(q).numpy()
        assert rel_l2(part.cpu().numpy(), ref) <= FP64_TOL


@pytest.mark.parametrize("name,nrhs", [("laplace_2d", 64), ("laplace_3d", 20), ("laplace_dl", 16), ("laplace_2d", 18),
                                       ("laplace_3d", 100), ("laplace_adl", 16)])
def test_mfma_multi_rhs_precision_form(name, nrhs):
    """FP32 matrix-core near field (R >= 16) against the FP64 apply on the same build. Widths
    that are multiples of 4 stage the charge tiles with LDS-direct loads (20 and 100: a partial
    64-channel block; 100: two blocks); 18 takes the register-staged kernel."""
    from skelfmm_amd.batched import SkelFMMBatched

    rng = np.random.default_rng(1)
    if name == "laplace_2d":
        X, op, g, leaf = annulus_points(6000, 5), ops.LAPLACE_2D, None, 64
    elif name == "laplace_3d":
        X, op, g, leaf = rng.random((12000, 3)), ops.LAPLACE_3D, None, 150
    else:
        # volume cloud with random unit normals (a surface with normals = positions is ill-conditioned
        # in FP32 coordinates; see test_fp32_double_layer_surface_conditioning)
        X = rng.random((12000, 3))
        g = rng.normal(size=(12000, 3))
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        op = ops.LAPLACE_DOUBLE_LAYER_3D if name == "laplace_dl" else ops.LAPLACE_ADJOINT_DOUBLE_LAYER_3D
        leaf = 150
    fmm = make_build(X, op, tol=1e-6, leaf_size=leaf, geometry=g)
    Q = rng.normal(size=(X.shape[0], nrhs))
    d64 = SkelFMMBatched(fmm.tree, op).build_from_copy(fmm)
    y64 = d64.matvec(torch.as_tensor(Q, device=_gpu())).cpu().numpy()
    d64.set_mfma_min_channels(-1)
    y64_valu = d64.matvec(torch.as_tensor(Q, device=_gpu())).cpu().numpy()
    assert rel_l2(y64, y64_valu) < 1e-13
    d32 = SkelFMMBatched(fmm.tree, op).build_from_copy(fmm, use_single_precision=True)
    q32 = torch.as_tensor(Q.astype(np.float32), device=_gpu())
    y_mfma = d32.matvec(q32).cpu().numpy().astype(np.float64)
    d32.set_mfma_min_channels(-1)
    y_valu = d32.matvec(q32).cpu().numpy().astype(np.float64)
    e_mfma = rel_l2(y_mfma, y64)
    e_valu = rel_l2(y_valu, y64)
    print(f"{name} R={nrhs}: MFMA {e_mfma:.2e} VALU {e_valu:.2e}", flush=True)
    bound = 2e-5 if name in ("laplace_dl", "laplace_adl") else 2e-6  # 1/r^2 double-layer sums cancel more
    assert e_mfma < bound and e_valu < bound
    assert rel_l2(y_mfma, y_valu) < 1e-6
    assert rel_l2(y64, BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(Q).numpy()) <= FP64_TOL


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("nrhs", [1, 2])
def test_graph_replay_matches_eager(single, nrhs):
    """Replays of a captured apply equal the eager apply; 1 RHS runs the symmetric near field
    (zero fill of the reverse-sum buffer, pairs of lists, fold), 2 RHS the one-sided kernel."""
    from skelfmm_amd.batched import SkelFMMBatched

    X = np.random.default_rng(3).random((20000, 3))
    fmm = make_build(X, ops.LAPLACE_3D, tol=1e-6, leaf_size=150)
    dev = SkelFMMBatched(fmm.tree, ops.LAPLACE_3D).build_from_copy(fmm, use_single_precision=single)
    assert nrhs > 1 or dev.use_symmetric
    dt = torch.float32 if single else torch.float64
    x = torch.randn(X.shape[0], nrhs, dtype=dt, device=_gpu()).squeeze(-1)
    out = torch.empty_like(x)
    eager = dev.apply_device(x).clone()
    g = dev.capture(x, out)
    for rep in range(4):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager) or rel_l2(out.cpu().numpy(), eager.cpu().numpy()) < (1e-6 if single else 1e-14)
        # eager work between replays
        tmp = dev.apply_device(torch.randn_like(x))
        torch.cuda.synchronize()
        del tmp
    x.copy_(x * 2.0)
    g.replay()
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), 2.0 * eager.cpu().numpy()) < (1e-6 if single else 1e-14)


def test_captured_graph_outlives_setters():
    """A captured graph addresses the plan's work vectors and the layout's item tables. The
    setters replace those; the graph must keep its own copies alive (replays stay correct and
    write nothing into memory handed out since), report itself stale, and the solver operator
    must capture again and then follow the new settings."""
    from skelfmm_amd.batched import SkelFMMBatched
    from skelfmm_amd.solvers import FMMOperator

    X = np.random.default_rng(4).random((20000, 3))
    fmm = make_build(X, ops.LAPLACE_3D, tol=1e-6, leaf_size=150)
    dev = SkelFMMBatched(fmm.tree, ops.LAPLACE_3D).build_from_copy(fmm, use_single_precision=True)
    assert dev.use_symmetric
    x = torch.randn(X.shape[0], dtype=torch.float32, device=_gpu())
    out = torch.empty_like(x)
    eager = dev.apply_device(x).clone()
    g = dev.capture(x, out)
    A = FMMOperator(dev, graph=True)
    y_graph = A(x)
    gen0 = dev.generation
    dev.set_near_field_split(1024)
    dev.set_symmetric(False)
    assert dev.generation > gen0 and g.stale
    # memory freed by the setters would be handed out again here
    fills = [torch.full((2 * dev.nslots,), 7.0, dtype=torch.float32, device=_gpu()) for _ in range(6)]
    fills += [torch.full((int(n),), 3, dtype=torch.int32, device=_gpu()) for n in (dev._nf["items"].numel(),) * 4]
    for _ in range(3):
        out.zero_()
        g.replay()
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), eager.cpu().numpy()) < 1e-6
    for f in fills:
        assert bool((f == f.reshape(-1)[0]).all())
    # the solver operator notices the change and captures again
    y_new = A(x)
    assert not A._graphs[(X.shape[0],)][2].stale
    ref = dev.apply_device(x)
    torch.cuda.synchronize()
    assert rel_l2(y_new.cpu().numpy(), ref.cpu().numpy()) < 1e-6
    assert rel_l2(y_graph.cpu().numpy(), eager.cpu().numpy()) < 1e-6


def test_fp32_double_layer_surface_conditioning():
    """Double layer on the unit sphere with normals = positions: rn = (x - y).n = -r^2/2 is
    ill-conditioned in the coordinates (rounding only the points to FP32 costs ~4e-3 in an exact
    sum). The precision form forms rn in FP64 from FP64 rows (absolute points, normals, n.x), so
    it must beat the oracle's restatement of the reference's single-precision form on the same
    build by far, and the FP32 matrix-core and VALU paths (same rn, box-centre shift rounded at
    different places) must agree to FP32 rounding."""
    from skelfmm_amd.batched import SkelFMMBatched

    X = sphere_points(5000, 4)
    fmm = make_build(X, ops.LAPLACE_DOUBLE_LAYER_3D, tol=1e-6, leaf_size=100, geometry=X)
    Q = np.random.default_rng(1).normal(size=(X.shape[0], 16))
    y64 = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(Q).numpy()
    ref32 = BatchedApplyCPU(fmm, use_single_precision=True, bytes_budget=2e8).matvec(Q.astype(np.float32)).numpy()
    d32 = SkelFMMBatched(fmm.tree, fmm.operator).build_from_copy(fmm, use_single_precision=True)
    q32 = torch.as_tensor(Q.astype(np.float32), device=_gpu())
    y_mfma = d32.matvec(q32).cpu().numpy().astype(np.float64)
    d32.set_mfma_min_channels(-1)
    y_valu = d32.matvec(q32).cpu().numpy().astype(np.float64)
    e_ref = rel_l2(ref32.astype(np.float64), y64)
    e_gpu = rel_l2(y_mfma, y64)
    print(f"sphere DL FP32: oracle single-precision form {e_ref:.2e}, GPU precision form {e_gpu:.2e}", flush=True)
    e_paths = rel_l2(y_mfma, y_valu)
    print(f"  matrix-core vs VALU near field: {e_paths:.2e}", flush=True)
    # what remains is the FP32 rounding of the box-centred distances (1/r^3 amplifies it on the
    # closest pairs of this coarse 5,000-point surface)
    assert e_gpu <= 0.01 * e_ref and e_gpu <= 2e-5
    assert e_paths <= 2e-5


def test_autotune_picks_setting_and_keeps_parity(tmp_path, monkeypatch):
    from skelfmm_amd import autotune
    from skelfmm_amd.batched import SkelFMMBatched

    monkeypatch.setenv("SKELFMM_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    X = np.random.default_rng(4).random((15000, 3))
    fmm = make_build(X, ops.LAPLACE_3D, tol=1e-6, leaf_size=150)
    dev = SkelFMMBatched(fmm.tree, ops.LAPLACE_3D).build_from_copy(fmm)
    best = autotune.tune(dev, 1, reps=1, splits=(2048, 16384))
    assert best["max_sources"] in (2048, 16384) and not best["cached"]
    assert dev.layout.max_sources == best["max_sources"]
    q = np.random.RandomState(0).randn(X.shape[0])
    y = dev.matvec(torch.as_tensor(q, device=_gpu())).cpu().numpy()
    ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(q).numpy()
    assert rel_l2(y, ref) <= FP64_TOL
    assert best["sym_unroll"] in (1, 2, 4) and dev.sym_unroll == best["sym_unroll"]
    assert best["item_order"] in ("heavy", "xcd") and dev.item_order == best["item_order"]
    again = autotune.tune(dev, 1, reps=1, splits=(2048, 16384))
    assert again["cached"] and again["max_sources"] == best["max_sources"]


@pytest.mark.parametrize("name,single,nrhs", [("laplace_3d_cube", False, 1), ("laplace_3d_cube", True, 1),
                                               ("laplace_2d_annulus", False, 20), ("laplace_2d_annulus", True, 20),
                                               ("helmholtz_dl_sphere", False, 1), ("helmholtz_3d_cube", True, 1)])
def test_item_orders_keep_parity(name, single, nrhs):
    """The XCD launch order of the work items (spatial runs per XCD, padded with empty items) against
    heaviest first: the same sums in another atomic order (FP64 within 1e-14, FP32 within 1e-6), and
    FP64 within 1e-12 of the oracle; covers the symmetric, one-sided and matrix-core kernels."""
    from skelfmm_amd.batched import SkelFMMBatched

    X, op, kappa, geom, leaf = _case(name)
    print(f"case {name} single={single} nrhs={nrhs}", flush=True)
    fmm = make_build(X, op, kappa=kappa, tol=1e-6, leaf_size=leaf, geometry=geom)
    dev = SkelFMMBatched(fmm.tree, op, kappa=kappa).build_from_copy(fmm, use_single_precision=single)
    rs = np.random.RandomState(7)
    q = rs.randn(X.shape[0], nrhs) if nrhs > 1 else rs.randn(X.shape[0])
    if op.is_complex:
        q = q + 1j * rs.randn(*q.shape)
    if single and not op.is_complex:
        q = q.astype(np.float32)
    x = torch.as_tensor(q).to(_gpu())
    y_heavy = dev.matvec(x).cpu().numpy()
    for ms in (1024, 16384):
        dev.set_near_field_split(ms)
        dev.set_item_order("xcd")
        assert dev.item_order == "xcd"
        nitems = int(dev._nf["items"].shape[0] // 4)
        assert nitems >= dev.layout.items.shape[0]  # padding items included
        y_xcd = dev.matvec(x).cpu().numpy()
        assert rel_l2(y_xcd, y_heavy) <= (1e-6 if single else 1e-14)
        dev.set_item_order("heavy")
    if not single:
        ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(q).numpy()
        assert rel_l2(y_xcd, ref) <= FP64_TOL
    with pytest.raises(ValueError):
        dev.set_item_order("random")


@pytest.mark.parametrize("name", ["laplace_2d_annulus", "laplace_3d_cube", "helmholtz_3d_cube", "helmholtz_2d_square",
                                  "laplace_dl_sphere", "laplace_adl_sphere", "helmholtz_dl_sphere",
                                  "helmholtz_adl_sphere"])
@pytest.mark.parametrize("single", [False, True])
def test_symmetric_near_field_matches_one_sided(name, single):
    """Pairs evaluated once for both directions (p2p_sym_kernel) against the one-sided kernel:
    FP64 to 1e-13 (only the summation order differs), FP32 form to 3e-6 of the result. Double
    layers share g(r) between the directions and take each direction's normal projection."""
    from skelfmm_amd.batched import SkelFMMBatched

    X, op, kappa, geom, leaf = _case(name)
    print(f"case {name} single={single} N={X.shape[0]}", flush=True)
    fmm = make_build(X, op, kappa=kappa, tol=1e-6, leaf_size=leaf, geometry=geom)
    dev = SkelFMMBatched(fmm.tree, op, kappa=kappa).build_from_copy(fmm, use_single_precision=single)
    assert dev.use_symmetric and dev.evaluations < dev.interactions
    rs = np.random.RandomState(2)
    q = rs.randn(X.shape[0])
    if op.is_complex:
        q = q + 1j * rs.randn(X.shape[0])
    x = torch.as_tensor(q).to(_gpu())
    y_sym = dev.matvec(x).cpu().numpy()
    dev.set_symmetric(False)
    y_one = dev.matvec(x).cpu().numpy()
    dev.set_symmetric(True)
    y_again = dev.matvec(x).cpu().numpy()
    err = rel_l2(y_sym, y_one)
    print(f"symmetric vs one-sided: {err:.2e}")
    assert err <= (3e-6 if single else 1e-13)
    assert rel_l2(y_again, y_sym) <= (1e-6 if single else 1e-14)  # atomics reorder sums only
    for un in (1, 4):  # inner-loop unrolls of the symmetric kernel give the same sums
        dev.set_sym_unroll(un)
        assert rel_l2(dev.matvec(x).cpu().numpy(), y_sym) <= (1e-6 if single else 1e-14)
    dev.set_sym_unroll(0)
    if not single:
        ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(q).numpy()
        if op is not ops.HELMHOLTZ_2D:
            assert rel_l2(y_sym, ref) <= FP64_TOL


@pytest.mark.parametrize("name,block", [("laplace_3d_cube", 64), ("laplace_3d_cube", 32), ("helmholtz_2d_square", 32),
                                        ("helmholtz_dl_sphere", 64), ("laplace_adl_sphere", 48)])
@pytest.mark.parametrize("single", [False, True])
def test_self_blocks_match_one_sided_self(name, block, single):
    """Self segments evaluated symmetrically in blocks (pairs of different blocks in the symmetric
    kernel, diagonal blocks one-sided) against one-sided self segments: FP64 to 1e-13, FP32 form
    to 3e-6; back to block 0 gives the first result again; the autotuner's self-block stage
    keeps parity."""
    from skelfmm_amd import autotune
    from skelfmm_amd.batched import SkelFMMBatched

    X, op, kappa, geom, leaf = _case(name)
    print(f"case {name} block={block} single={single} N={X.shape[0]}", flush=True)
    fmm = make_build(X, op, kappa=kappa, tol=1e-6, leaf_size=leaf, geometry=geom)
    dev = SkelFMMBatched(fmm.tree, op, kappa=kappa).build_from_copy(fmm, use_single_precision=single)
    rs = np.random.RandomState(2)
    q = rs.randn(X.shape[0])
    if op.is_complex:
        q = q + 1j * rs.randn(X.shape[0])
    x = torch.as_tensor(q).to(_gpu())
    y0 = dev.matvec(x).cpu().numpy()
    pairs0, ev0 = dev.sym_pairs, dev.evaluations
    dev.set_self_block(block)
    assert dev.self_block == block and dev.sym_pairs > pairs0 and dev.evaluations < ev0
    assert dev.one_interactions + 2 * dev.sym_pairs == dev.interactions
    y1 = dev.matvec(x).cpu().numpy()
    err = rel_l2(y1, y0)
    print(f"self blocks of {block} vs one-sided self segments: {err:.2e}")
    assert err <= (3e-6 if single else 1e-13)
    dev.set_near_field_split(1024)  # shorter source runs over the same tables
    assert rel_l2(dev.matvec(x).cpu().numpy(), y1) <= (1e-6 if single else 1e-14)
    dev.set_self_block(0)
    assert dev.sym_pairs == pairs0
    assert rel_l2(dev.matvec(x).cpu().numpy(), y0) <= (1e-6 if single else 1e-14)
    if name == "laplace_3d_cube" and block == 64:
        best = autotune.tune(dev, 1, use_cache=False, splits=(4096,), levels=False)
        assert best["self_block"] in (0,) + tuple(autotune.SELF_BLOCKS) and dev.self_block == best["self_block"]
        assert rel_l2(dev.matvec(x).cpu().numpy(), y0) <= (3e-6 if single else 1e-13)
    with pytest.raises(ValueError):
        dev.set_self_block(-1)


@pytest.mark.parametrize("single,nrhs", [(False, 1), (True, 1), (False, 1j), (True, 20)])
def test_level_variants_keep_parity(single, nrhs):
    """Every launch variant of the interpolation stages (loads in flight for one channel, matrix
    cores or VALU GEMM at >= 16 FP32 channels) gives the same apply: FP64 within 1e-12 of the
    oracle, FP32 variants within FP32 rounding of each other; the level autotuner keeps parity."""
    from skelfmm_amd import autotune
    from skelfmm_amd.batched import SkelFMMBatched

    rng = np.random.default_rng(8)
    if nrhs == 1j:  # complex vectors (Helmholtz), one channel
        X, op, kappa, R = rng.random((8000, 3)), ops.HELMHOLTZ_3D, 4.0, 1
    else:
        X, op, kappa, R = rng.random((12000, 3)), ops.LAPLACE_3D, 0.0, int(nrhs)
    fmm = make_build(X, op, kappa=kappa, tol=1e-6, leaf_size=150)
    dev = SkelFMMBatched(fmm.tree, op, kappa=kappa).build_from_copy(fmm, use_single_precision=single)
    q = rng.normal(size=(X.shape[0], R)) if R > 1 else rng.normal(size=X.shape[0])
    if op.is_complex:
        q = q + 1j * rng.normal(size=q.shape)
    x = dev.prepare_input(torch.as_tensor(q).to(_gpu()))
    width = (2 if dev.complex_vectors else 1) * R
    choices = autotune.level_variant_choices(dev, width)
    ys = {}
    for v in choices:
        dev.set_level_variants(width, up=[v] * len(dev.layout.up), down=[v] * len(dev.layout.down))
        ys[v] = dev.apply_device(x).cpu().numpy()
    print(f"single={single} width={width}: variants {choices}", flush=True)
    if not single:
        ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(q).numpy()
        for v, y in ys.items():
            assert rel_l2(y, ref) <= FP64_TOL, v
    for v, y in ys.items():
        assert rel_l2(y, ys[0]) <= (1e-6 if single else 1e-14), v
    res = autotune.tune_levels(dev, width, reps=1)
    assert len(res["up"]) == len(dev.layout.up) and len(res["down"]) == len(dev.layout.down)
    assert dev.level_variants(width) == {"up": res["up"], "down": res["down"]}
    y = dev.apply_device(x).cpu().numpy()
    assert rel_l2(y, ys[0]) <= (1e-6 if single else 1e-14)


@pytest.mark.parametrize("op_name", ["laplace_3d", "laplace_2d"])
def test_coincident_points_in_both_forms(op_name):
    """Two distinct points at the same place, every other charge zero, one box (all direct): the
    FP64 form follows the reference's clamp r^2 >= 1e-15 (util_batched.py:175-185); the FP32 form
    evaluates K at r^2 = 1e-20 (its r^2 is accumulated on a 1e-20 bias: 1/r = 1e10 in 3D, 0.5 log
    r^2 = -23.03 in 2D) rather than the reference FP32 clamp r^2 >= 1e-8. Self pairs stay zero."""
    op = ops.LAPLACE_3D if op_name == "laplace_3d" else ops.LAPLACE_2D
    d = 3 if op is ops.LAPLACE_3D else 2
    X = np.random.default_rng(11).random((400, d))
    X[7] = X[123]
    q = np.zeros(X.shape[0])
    q[7] = 2.0
    fmm = make_build(X, op, tol=1e-6, leaf_size=500)
    y64, _ = _device_apply(fmm, q)
    y32, _ = _device_apply(fmm, q.astype(np.float32), single=True)
    k64 = 1.0 / np.sqrt(1e-15) if d == 3 else 0.5 * np.log(1e-15)
    k32 = 1.0 / np.sqrt(1e-20) if d == 3 else 0.5 * np.log(1e-20)
    print(f"{op_name}: coincident pair FP64 {y64[123] / 2:.6e} FP32 {float(y32[123]) / 2:.6e}", flush=True)
    assert abs(y64[123] / 2.0 - k64) <= 1e-12 * abs(k64)
    assert abs(float(y32[123]) / 2.0 - k32) <= 1e-6 * abs(k32)
    assert y64[7] == 0.0 and float(y32[7]) == 0.0  # the charged point's own self pair is zero


def test_torch_operator_matches_c_abi_and_checks_arguments():
    """torch.ops.skelfmm.apply (the PyTorch extension) and skf_apply through ctypes on the same
    plan give the same FP64 apply (atomic order aside), and the operator refuses mismatched
    tensors with the reason."""
    from skelfmm_amd import hip
    from skelfmm_amd.batched import SkelFMMBatched

    X = np.random.default_rng(9).random((3000, 2))
    fmm = make_build(X, ops.LAPLACE_2D, tol=1e-6, leaf_size=64)
    dev = SkelFMMBatched(fmm.tree, ops.LAPLACE_2D).build_from_copy(fmm)
    x = torch.as_tensor(np.random.RandomState(1).randn(X.shape[0]), device=_gpu())
    p = dev._plan(1)
    y_op = torch.empty_like(x)
    hip.torch_ops().apply(p["handle"].value, x, y_op)
    y_c = torch.empty_like(x)
    hip.check(hip.lib().skf_apply(p["handle"], hip.ptr(x), hip.ptr(y_c), hip.stream_handle(x.device)))
    torch.cuda.synchronize()
    assert rel_l2(y_op.cpu().numpy(), y_c.cpu().numpy()) <= 1e-14
    ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(x.cpu().numpy()).numpy()
    assert rel_l2(y_op.cpu().numpy(), ref) <= FP64_TOL
    with pytest.raises(RuntimeError, match="float32 or float64"):
        hip.torch_ops().apply(p["handle"].value, x, torch.empty(x.shape, dtype=torch.float32, device=x.device))
    with pytest.raises(RuntimeError, match="same size"):
        hip.torch_ops().apply(p["handle"].value, x, torch.empty(10, dtype=x.dtype, device=x.device))
    with pytest.raises(RuntimeError, match="null plan"):
        hip.torch_ops().apply(0, x, y_op)


@pytest.mark.parametrize("single", [False, True])
def test_mfma_store_rows_and_atomic_sums_agree(single):
    """Matrix-core near field with single-writer rows (one work item per target block: the sums
    are stored, source parts combined in LDS) against split source runs (several items per block:
    atomic sums): the same apply up to summation order; FP64 also against the oracle."""
    from skelfmm_amd.batched import SkelFMMBatched

    X = annulus_points(6000, 5)
    fmm = make_build(X, ops.LAPLACE_2D, tol=1e-6, leaf_size=64)
    dev = SkelFMMBatched(fmm.tree, ops.LAPLACE_2D).build_from_copy(fmm, use_single_precision=single)
    Q = np.random.RandomState(11).randn(X.shape[0], 24)
    x = torch.as_tensor(Q.astype(np.float32) if single else Q, device=_gpu())
    dev.set_near_field_split(16384)
    assert dev.rows_exclusive
    y_store = dev.matvec(x).cpu().numpy()
    dev.set_near_field_split(64)
    assert not dev.rows_exclusive
    y_atomic = dev.matvec(x).cpu().numpy()
    assert rel_l2(y_store, y_atomic) <= (1e-6 if single else 1e-14)
    if not single:
        ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(Q).numpy()
        assert rel_l2(y_store, ref) <= FP64_TOL
        assert rel_l2(y_atomic, ref) <= FP64_TOL
