"""GPU tree and list construction against the host statement, bit for bit.

``GPUBalancedTree`` (HIP kernels) and ``BalancedTree`` (numpy; itself pinned by brute-force
neighbour checks in tests/test_tree.py) must agree on every integer: the sorted point order,
level boundaries, box codes, point ranges, integer coordinates, children, parents, leaves, the
colleague and coarse neighbour rows of every box and the leaf / tree / coarse pair lists of a
build. Clouds: uniform 2D and 3D, a Gaussian mixture (deep, heavily balanced levels), the unit
sphere, a curvy annulus, coincident points, and the smallest inputs.
"""

import numpy as np
import pytest
import torch

from conftest import annulus_points, make_build, sphere_points
from skelfmm_amd import operators as ops
from skelfmm_amd.tree import BalancedTree

pytestmark = pytest.mark.gpu


def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")


def _mixture(n, seed=3):
    rng = np.random.default_rng(seed)
    c = rng.random((16, 3))
    w = rng.dirichlet(np.ones(16))
    comp = rng.choice(16, size=n, p=w)
    return np.clip(c[comp] + 0.02 * rng.standard_normal((n, 3)), 0.0, 1.0)


def _clouds():
    rng = np.random.default_rng(0)
    yield "uniform3d", rng.random((60000, 3)), 100
    yield "uniform2d", rng.random((40000, 2)), 64
    yield "mixture3d", _mixture(120000), 64
    yield "sphere", sphere_points(50000, 1), 64
    yield "annulus", annulus_points(30000, 2), 40
    dup = rng.random((3000, 3))
    dup[1000:1500] = dup[7]  # coincident points
    yield "coincident", dup, 600
    yield "single_point", np.array([[0.3, 0.7, 0.1]]), 5
    yield "below_leaf", rng.random((17, 2)), 64
    yield "leaf_plus_one", rng.random((65, 3)), 64


CLOUDS = list(_clouds())

TABLES = ("_perm", "_codes", "level_sep", "_box_level", "_box_code", "_box_start", "_box_end", "_box_icoord",
          "_children", "_parent", "_is_leaf", "_leaves")


@pytest.mark.parametrize("case", CLOUDS, ids=lambda c: c[0])
def test_gpu_tree_equals_host_tree(case):
    from skelfmm_amd.gpu_tree import GPUBalancedTree

    _gpu()
    name, X, leaf = case
    host = BalancedTree(X, leaf_size=leaf)
    dev = GPUBalancedTree(X, leaf_size=leaf)
    print(f"{name}: N={X.shape[0]} leaf={leaf} levels={host.nlevels} boxes={host.nboxes} leaves={host.nleaves}",
          flush=True)
    assert (dev.nlevels, dev.nboxes, dev.nleaves) == (host.nlevels, host.nboxes, host.nleaves)
    assert dev.L0 == host.L0 and np.array_equal(dev.c0, host.c0)
    for k in TABLES:
        a, b = getattr(dev, k), getattr(host, k)
        assert a.shape == b.shape and np.array_equal(a, b), k
    boxes = np.arange(host.nboxes)
    assert np.array_equal(dev.get_boxes_colleague_neigh_fullvec(boxes), host.get_boxes_colleague_neigh_fullvec(boxes))
    assert np.array_equal(dev.get_boxes_coarse_neigh_fullvec(boxes), host.get_boxes_coarse_neigh_fullvec(boxes))
    for lev in range(host.nlevels):
        assert np.array_equal(dev.get_leaves_above(lev), host.get_leaves_above(lev))


@pytest.mark.parametrize("case", [c for c in CLOUDS if c[0] in ("uniform3d", "mixture3d", "annulus")],
                         ids=lambda c: c[0])
def test_gpu_pair_lists_equal_host_lists(case):
    """A build on the GPU tree forms the five pair lists from the device rows; they must equal the
    host build's lists entry for entry (same order), and so must the box layout."""
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.gpu_tree import GPUBalancedTree
    from skelfmm_amd.layout import build_layout

    _gpu()
    name, X, leaf = case
    op = ops.LAPLACE_3D if X.shape[1] == 3 else ops.LAPLACE_2D
    host = make_build(X, op, tol=1e-4, leaf_size=leaf)
    gt = GPUBalancedTree(X, leaf_size=leaf)
    dev = RecursiveSkeletonization(gt, op, kappa=0.0, tol=1e-4)
    dev.skel_tree(npoints_max=500)
    dev.setup_lists()
    assert dev.root_level == host.root_level
    for k in ("leaf_uskel_from_qskel_list", "tree_uskel_from_qskel_list", "uskel_from_qskel_list",
              "uskel_from_q_list", "u_from_qskel_list", "u_from_q_list", "root_boxes", "skel_leaves"):
        a, b = getattr(dev, k), getattr(host, k)
        assert a.shape == b.shape and np.array_equal(a, b), k
    assert np.array_equal(dev.rank_list, host.rank_list)
    La, Lb = build_layout(dev), build_layout(host)
    for k in ("off", "slot_point", "leaf_slot", "leaf_user", "items", "g_pt", "g_n", "e_pt", "e_n", "e_q", "e_flags"):
        assert np.array_equal(getattr(La, k), getattr(Lb, k)), k
    print(f"{name}: {dev.leaf_uskel_from_qskel_list.shape[0]} leaf pairs, {dev.tree_uskel_from_qskel_list.shape[0]} "
          f"tree pairs, {dev.uskel_from_q_list.shape[0]} coarse pairs identical", flush=True)


def test_gpu_tree_rejects_bad_input():
    from skelfmm_amd.gpu_tree import GPUBalancedTree

    _gpu()
    with pytest.raises(ValueError):
        GPUBalancedTree(np.zeros((0, 3)))
    with pytest.raises(ValueError):
        GPUBalancedTree(np.zeros((10, 4)))
    with pytest.raises(ValueError):
        GPUBalancedTree(np.random.default_rng(0).random((10, 3)), leaf_size=0)
    with pytest.raises(ValueError, match="coincident"):
        GPUBalancedTree(np.zeros((40, 2)), leaf_size=8)  # more coincident points than a leaf holds


@pytest.mark.parametrize("name", ["laplace_3d", "helmholtz_3d", "helmholtz_double_layer_3d", "laplace_2d"])
def test_device_near_field_tables_equal_host_tables(name):
    """The near-field work tables formed on the device (rows, symmetric split, groups, entries,
    items heaviest first) equal the host tables entry for entry."""
    from skelfmm_amd.layout import NearLists, build_layout

    _gpu()
    rng = np.random.default_rng(5)
    if name == "laplace_2d":
        X, op, kappa, g = annulus_points(20000, 3), ops.LAPLACE_2D, 0.0, None
    elif name == "helmholtz_double_layer_3d":
        X = sphere_points(20000, 2)
        op, kappa, g = ops.HELMHOLTZ_DOUBLE_LAYER_3D, 6.0, X
    else:
        X = _mixture(40000, seed=7)
        op, kappa, g = (ops.LAPLACE_3D, 0.0, None) if name == "laplace_3d" else (ops.HELMHOLTZ_3D, 8.0, None)
    f = make_build(X, op, kappa=kappa, tol=1e-5, leaf_size=64, geometry=g)
    for ms in (512, 16384):
        host = build_layout(f, max_sources=ms)
        dev = build_layout(f, max_sources=ms, near_field_device=torch.device("cuda", 0))
        fields = NearLists.FIELDS + ("items", "item_cost")
        for k in fields + ("interactions", "evaluations"):
            a, b = getattr(dev, k), getattr(host, k)
            assert np.array_equal(np.asarray(a), np.asarray(b)), k
        assert (dev.split is None) == (host.split is None)
        for P, Q in zip(dev.split or (), host.split or ()):
            assert P.tpi == Q.tpi and P.interactions == Q.interactions
            for k in fields:
                assert np.array_equal(getattr(P, k), getattr(Q, k)), "split." + k
    print(f"{name}: {host.items.shape[0]} items, split {host.split is not None}: device tables identical", flush=True)
