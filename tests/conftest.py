import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def hip_lib():
    """The in-tree HIP library, compiled on demand when missing or stale (no GPU needed to load it)."""
    from skelfmm_amd import hip

    return hip.lib()


def make_build(points, operator, kappa=0.0, tol=1e-6, leaf_size=64, geometry=None, npoints_max=500):
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.tree import BalancedTree

    tree = BalancedTree(points, leaf_size=leaf_size)
    fmm = RecursiveSkeletonization(tree, operator, kappa=kappa, tol=tol, geometry=geometry)
    fmm.skel_tree(npoints_max=npoints_max)
    fmm.setup_lists()
    return fmm


def sphere_points(n, seed=0):
    v = np.random.default_rng(seed).normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def annulus_points(n, seed=0):
    """Curvy annulus: r = rho(theta) (0.5 + 0.5 u), rho = 1 + 0.25 cos(6 theta)."""
    rng = np.random.default_rng(seed)
    th = rng.random(n) * 2 * np.pi
    u = rng.random(n)
    r = (1 + 0.25 * np.cos(6 * th)) * (0.5 + 0.5 * u)
    return np.stack([r * np.cos(th), r * np.sin(th)], axis=1)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))
