"""User-defined operators: the expression language, its numpy / torch evaluators, the
run-time compiled near field.

CPU: the validator rejects everything outside the language; formulas equal to built-in
kernels reproduce the built-in dense kernels bit for bit (``rinv`` = Laplace 3D, ``0.5 *
log(r2)`` = Laplace 2D) and so give the same host build; a Gaussian matches its closed
form; ``torch_apply`` matches the dense matrix; the generated HIP source compiles with
hiprtc for both precisions (compile only, no device).

GPU: the apply of a formula equal to a built-in kernel equals the built-in apply on the
same build (FP64 to 1e-12, FP32 form to 2e-6), over 1, 3 and 8 channels (channel blocks 1,
4 and 8); a screened Coulomb kernel (3D) and a dipole kernel (2D, not symmetric) against
the direct sum.
"""

import numpy as np
import pytest
import torch

from conftest import make_build, rel_l2
from skelfmm_amd import dense, operators as ops
from skelfmm_amd.jit import UserKernel, parse_expression


@pytest.mark.parametrize("bad", [
    "r.real", "__import__('os')", "lambda: 1", "r2[0]", "foo", "'1'", "r2 < 1", "True", "exp(x=r2)",
    "unknown(r2)", "exp", "r2 if r2 else 1", "min(r2)", "exp(r2, 2)", "r2 % 2", "", "  ", "(r2,)", "r2 @ r2",
])
def test_expression_validator_rejects(bad):
    with pytest.raises(ValueError):
        parse_expression(bad)


def test_expression_validator_accepts_the_language():
    for ok in ["rinv", "1/r", "exp(-param * r2)", "r2 * log(r2)", "dx * rinv**3", "-(dx + dy + dz) / r**2",
               "sqrt(r2) + min(r, 2.5) - max(r, 1e-3)", "tanh(r) * cos(param * r) + erf(r) + abs(dx)", "pi * r2"]:
        parse_expression(ok)


def _pts(n, d, seed=0):
    X = np.random.default_rng(seed).random((n, d))
    X[5] = X[3]  # a coincident pair
    return X


def test_formulas_equal_to_builtin_kernels_are_bit_identical():
    X, Y = _pts(300, 3), _pts(200, 3, 1)
    lap3 = ops.custom_operator("my_laplace_3d", "rinv", ndim=3)
    np.testing.assert_array_equal(lap3.numpy_matrix(X, X), dense.laplace_3d(X, X))
    np.testing.assert_array_equal(lap3.numpy_matrix(X, Y, assume_not_self_interaction=True),
                                  dense.laplace_3d(X, Y, assume_not_self_interaction=True))
    X2 = _pts(300, 2)
    lap2 = ops.custom_operator("my_laplace_2d", "0.5 * log(r2)", ndim=2)
    np.testing.assert_array_equal(lap2.numpy_matrix(X2, X2), dense.laplace_2d(X2, X2))
    assert lap3.is_symmetric and lap2.is_symmetric and lap3.kernel_id == 17 and lap2.kernel_id == 16


def test_gaussian_and_directional_closed_forms():
    X, Y = _pts(120, 3), _pts(90, 3, 1)
    g = ops.custom_operator("gauss", "exp(-param * r2)", ndim=3)
    D = X[:, None, :] - Y[None, :, :]
    r2 = np.sum(D * D, axis=-1)
    np.testing.assert_allclose(g.numpy_matrix(X, Y, param=7.0), np.exp(-7.0 * r2), rtol=1e-13)
    d = ops.custom_operator("dipole_x", "dx * rinv**3", ndim=3)
    assert not d.is_symmetric
    np.testing.assert_allclose(d.numpy_matrix(X, Y), D[..., 0] / r2 ** 1.5, rtol=1e-13)
    K = g.numpy_matrix(X, X, param=2.0)
    assert np.all(np.diag(K)[np.arange(120) != 5] == 0) and K[5, 3] == 0  # coincident points contribute 0


def test_torch_apply_matches_dense_matrix():
    X, Y = _pts(64, 3), _pts(48, 3, 1)
    g = ops.custom_operator("gauss", "exp(-param * r2) * (1 + r)", ndim=3)
    q = np.random.default_rng(2).normal(size=(48, 2))
    out = g.torch_apply(torch.as_tensor(q)[None], torch.as_tensor(X)[None], torch.as_tensor(Y)[None], param=3.0)
    np.testing.assert_allclose(out[0].numpy(), g.numpy_matrix(X, Y, param=3.0) @ q, rtol=1e-13, atol=1e-13)


def test_host_build_of_an_equivalent_formula_is_identical():
    X = np.random.default_rng(4).random((6000, 3))
    f0 = make_build(X, ops.LAPLACE_3D, tol=1e-6, leaf_size=100)
    f1 = make_build(X, ops.custom_operator("my_laplace_3d", "rinv", ndim=3), tol=1e-6, leaf_size=100)
    np.testing.assert_array_equal(f0.rank_list, f1.rank_list)
    for b in range(f0.nboxes):
        if f0.cand[b] is not None:
            np.testing.assert_array_equal(f0.cand[b], f1.cand[b])


def test_generated_source_compiles_for_gfx950(hip_lib):
    from skelfmm_amd.jit import compile_kernels

    for expr, d in [("exp(-param * r2) * rinv", 3), ("r2 * log(r2)", 2), ("tanh(r) + erf(dx) + r**1.5 + rinv**-2", 3)]:
        uk = UserKernel(expr, d)
        src = uk.source()
        assert "kernel_value<double, %d>" % uk.kernel_id in src
        for prec in (0, 1):
            assert compile_kernels(uk, prec, load=False) > 0


def test_generated_source_reports_compiler_errors(hip_lib):
    from skelfmm_amd.jit import compile_kernels

    uk = UserKernel("rinv", 3)
    uk.c_expression = "rinv +"  # a mistake planted after validation: hiprtc must refuse it
    with pytest.raises(RuntimeError, match="compilation failed"):
        compile_kernels(uk, 0, load=False)


def test_gpu_build_rejects_user_operators():
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.tree import BalancedTree

    X = np.random.default_rng(0).random((500, 3))
    f = RecursiveSkeletonization(BalancedTree(X, 50), ops.custom_operator("g", "rinv"), kappa=0.0, tol=1e-6)
    with pytest.raises(ValueError, match="host"):
        f.skel_tree(device="gpu")


# ------------------------------------------------------------------- GPU
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    return torch.device("cuda", 0)


def _apply(fmm, q, single):
    from skelfmm_amd.batched import SkelFMMBatched

    dev = SkelFMMBatched(fmm.tree, fmm.operator, kappa=fmm.kappa).build_from_copy(fmm, use_single_precision=single)
    y = dev.matvec(torch.as_tensor(q).to(_gpu()))
    torch.cuda.synchronize()
    y = y.cpu().numpy()
    return y if np.iscomplexobj(y) else y.astype(np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("nrhs", [1, 3, 8])
def test_user_formula_apply_equals_builtin(single, nrhs):
    X = np.random.default_rng(5).random((20000, 3))
    f0 = make_build(X, ops.LAPLACE_3D, tol=1e-6, leaf_size=150)
    f1 = make_build(X, ops.custom_operator("my_laplace_3d", "rinv", ndim=3), tol=1e-6, leaf_size=150)
    q = np.random.RandomState(1).randn(X.shape[0], nrhs).squeeze()
    y0 = _apply(f0, q, single)
    y1 = _apply(f1, q, single)
    e = rel_l2(y1, y0)
    print(f"single={single} nrhs={nrhs}: user vs built-in {e:.2e}", flush=True)
    assert e < (2e-6 if single else 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["yukawa3d", "dipole2d"])
def test_user_formula_apply_against_direct_sum(case):
    rng = np.random.default_rng(6)
    if case == "yukawa3d":  # screened Coulomb: Green's function of (Laplacian - param^2)
        X = rng.random((30000, 3))
        op, kappa, tol = ops.custom_operator("yukawa", "exp(-param * r) * rinv", ndim=3), 3.0, 1e-7
    else:  # 2D dipole d/dx log|x - y|: harmonic, not symmetric (stacked proxy block)
        X = rng.random((30000, 2))
        op, kappa, tol = ops.custom_operator("dipole2d", "dx / r2", ndim=2), 0.0, 1e-7
        assert not op.is_symmetric
    f = make_build(X, op, kappa=kappa, tol=tol, leaf_size=128)
    assert f.rank_list.sum() < f.bs_list.sum()  # the build compresses
    q = np.random.RandomState(1).randn(X.shape[0])
    y = _apply(f, q, single=False)
    assert not np.iscomplexobj(y)
    idx = np.random.default_rng(0).choice(X.shape[0], 300, replace=False)
    ref = op.numpy_matrix(X[idx], X, param=kappa) @ q
    e = rel_l2(y[idx], ref)
    print(f"{case}: rel-l2 vs direct sum {e:.2e}", flush=True)
    assert e < 10 * tol * max(1.0, np.log10(X.shape[0]))
