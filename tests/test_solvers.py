"""GMRES and the boundary-integral driver.

CPU (dense operators as torch CPU tensors): solutions against ``numpy.linalg.solve``; the
finite-termination property (a diagonalizable matrix with k distinct eigenvalues is solved
in at most k steps); restarts; columns solved in lockstep equal separate solves; monotone
residual history; the reported residual equals the true one; complex systems; and the
sphere double-layer discretization (dense) converging to the exact harmonic function, with
the self-cell correction reducing the error.

GPU: the same equation with the fast apply as the operator, FP64 and FP32 forms, eager and
graph-replayed products.
"""

import numpy as np
import pytest
import torch

from skelfmm_amd import operators as ops
from skelfmm_amd.solvers import fibonacci_sphere, gmres


def _dense(M):
    Mt = torch.as_tensor(M)
    return lambda x: Mt @ x


def _system(n, seed, cplx=False):
    rng = np.random.default_rng(seed)
    M = np.eye(n) + 0.4 * rng.normal(size=(n, n)) / np.sqrt(n)
    b = rng.normal(size=n)
    if cplx:
        M = M + 0.4j * rng.normal(size=(n, n)) / np.sqrt(n)
        b = b + 1j * rng.normal(size=n)
    return M, b


@pytest.mark.parametrize("cplx", [False, True])
def test_gmres_matches_direct_solve(cplx):
    M, b = _system(120, 0, cplx)
    res = gmres(_dense(M), torch.as_tensor(b), tol=1e-12, restart=120)
    assert res.converged.all()
    x = np.linalg.solve(M, b)
    assert np.linalg.norm(res.x.numpy() - x) / np.linalg.norm(x) < 1e-10
    true_rel = np.linalg.norm(b - M @ res.x.numpy()) / np.linalg.norm(b)
    assert true_rel <= 1e-12
    hist = np.array(res.residuals[0])
    assert np.all(np.diff(hist) <= 1e-15), "GMRES residuals do not increase within a cycle"
    assert abs(hist[-1] - true_rel) < 1e-11


def test_gmres_finite_termination():
    """k distinct eigenvalues -> the Krylov space holds the solution after k steps."""
    rng = np.random.default_rng(1)
    n, k = 200, 4
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    lam = np.repeat([1.0, 2.0, -3.0, 0.5], n // k)
    M = Q @ np.diag(lam) @ Q.T
    b = rng.normal(size=n)
    res = gmres(_dense(M), torch.as_tensor(b), tol=1e-11, restart=50)
    assert res.converged[0] and res.iterations[0] == k


def test_gmres_restarts_still_converge():
    M, b = _system(150, 2)
    full = gmres(_dense(M), torch.as_tensor(b), tol=1e-10, restart=150)
    short = gmres(_dense(M), torch.as_tensor(b), tol=1e-10, restart=5, maxiter=500)
    assert short.converged[0] and short.restarts > 1
    assert short.iterations[0] >= full.iterations[0]
    np.testing.assert_allclose(short.x.numpy(), full.x.numpy(), rtol=0, atol=1e-8)


def test_gmres_columns_in_lockstep_equal_separate_solves():
    M, _ = _system(100, 3)
    rng = np.random.default_rng(4)
    B = rng.normal(size=(100, 3))
    B[:, 1] *= 1e-3
    B[:, 2] = M @ np.ones(100)
    A = _dense(M)
    block = gmres(A, torch.as_tensor(B), tol=1e-11, restart=40)
    for r in range(3):
        one = gmres(A, torch.as_tensor(B[:, r].copy()), tol=1e-11, restart=40)
        assert one.iterations[0] == block.iterations[r]
        np.testing.assert_allclose(block.x[:, r].numpy(), one.x.numpy(), rtol=0,
                                   atol=1e-12 * np.abs(one.x.numpy()).max())


def test_gmres_zero_right_hand_side():
    M, b = _system(50, 5)
    B = np.stack([b, np.zeros(50)], axis=1)
    res = gmres(_dense(M), torch.as_tensor(B), tol=1e-12, restart=50)
    assert res.converged.all() and res.iterations[1] == 0
    assert np.all(res.x[:, 1].numpy() == 0)


def test_gmres_initial_guess():
    M, b = _system(80, 6)
    x = np.linalg.solve(M, b)
    res = gmres(_dense(M), torch.as_tensor(b), x0=torch.as_tensor(x), tol=1e-10)
    assert res.converged[0] and res.iterations[0] == 0


def _sphere_dense(n, correction):
    X, w = fibonacci_sphere(n)
    x0 = np.array([0.3, 2.0, -0.2])
    f = 1.0 / np.linalg.norm(X - x0, axis=1)
    K = ops.LAPLACE_DOUBLE_LAYER_3D.numpy_matrix(X, X, geom_src=X)
    np.fill_diagonal(K, 0.0)
    d = 0.5 + (np.sqrt(w / np.pi) / 4.0 if correction else 0.0)
    M = np.diag(np.broadcast_to(d, (n,))) + K * w[None, :]
    res = gmres(_dense(M), torch.as_tensor(f), tol=1e-12, restart=40)
    T = np.array([[0.1, 0.2, -0.3], [-0.4, 0.1, 0.2], [0.0, 0.0, 0.6], [0.5, -0.5, 0.1]])
    u = ops.LAPLACE_DOUBLE_LAYER_3D.numpy_matrix(T, X, geom_src=X) @ (w * res.x.numpy())
    exact = 1.0 / np.linalg.norm(T - x0, axis=1)
    return res, np.linalg.norm(u - exact) / np.linalg.norm(exact)


def test_sphere_double_layer_equation_dense():
    """Second kind: few iterations; the interior solution converges; the self-cell term helps."""
    r0, e0 = _sphere_dense(2000, correction=False)
    r1, e1 = _sphere_dense(2000, correction=True)
    assert r1.converged[0] and r1.iterations[0] < 25
    assert e1 < 2e-3 and e0 > 5 * e1
    _, e2 = _sphere_dense(500, correction=True)
    assert e2 > 1.5 * e1  # refinement reduces the error


def test_sphere_kernel_closed_form():
    """On the unit sphere with outward normals K(x, y) = 1 / (8 pi |x - y|)."""
    X, _ = fibonacci_sphere(300)
    K = ops.LAPLACE_DOUBLE_LAYER_3D.numpy_matrix(X, X, geom_src=X)
    D = np.linalg.norm(X[:, None] - X[None], axis=-1)
    off = ~np.eye(300, dtype=bool)
    np.testing.assert_allclose(K[off], 1.0 / (8 * np.pi * D[off]), rtol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("single", [False, True])
def test_sphere_dirichlet_with_fast_apply(single):
    from skelfmm_amd.solvers import laplace_sphere_dirichlet

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    kw = dict(tol=1e-6 if single else 1e-8, leaf_size=200, use_single_precision=single,
              gmres_tol=1e-5 if single else 1e-9)
    out = laplace_sphere_dirichlet(20000, **kw)
    print(f"single={single}: {out['iterations']} iterations, rel err {out['rel_err']:.2e}, "
          f"solve {out['solve_s']:.3f}s", flush=True)
    assert out["converged"] and out["iterations"] < 30
    assert out["rel_err"] < 1e-3
    eager = laplace_sphere_dirichlet(20000, graph=False, **kw)
    assert abs(eager["iterations"] - out["iterations"]) <= 1  # atomics order may move the last step
    assert np.linalg.norm(eager["sigma"] - out["sigma"]) / np.linalg.norm(out["sigma"]) < (1e-4 if single else 1e-8)


@pytest.mark.gpu
def test_fmm_operator_block_of_right_hand_sides():
    """Several right-hand sides share each multi-RHS apply and match separate solves."""
    from skelfmm_amd import SkelFMM
    from skelfmm_amd.solvers import FMMOperator

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    X, w = fibonacci_sphere(12000)
    fmm = SkelFMM(X, ops.LAPLACE_DOUBLE_LAYER_3D, geometry=X, tol=1e-8, leaf_size=200, multi_rhs=True)
    A = FMMOperator(fmm, weights=w, diag=0.5 + np.sqrt(w / np.pi) / 4.0, graph=True)
    rng = np.random.default_rng(0)
    B = torch.as_tensor(rng.normal(size=(12000, 3)), device=A.device)
    block = gmres(A, B, tol=1e-9, restart=30)
    assert block.converged.all()
    for r in range(3):
        one = gmres(A, B[:, r].contiguous(), tol=1e-9, restart=30)
        assert abs(int(one.iterations[0]) - int(block.iterations[r])) <= 1
        d = torch.linalg.vector_norm(one.x - block.x[:, r]) / torch.linalg.vector_norm(one.x)
        assert float(d) < 1e-8
