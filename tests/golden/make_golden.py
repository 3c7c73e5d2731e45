"""Write the golden fixtures of tests/golden (run from the repository root).

Each fixture holds seeded inputs (points, charges, normals), the FP64 direct
sum from the oracle's closed forms and the oracle's apply on this repository's
build, with the build parameters. The reference ships no fixtures of its own
(its tests draw points through the absent simpletree package), so these are
generated by the oracle; the direct sums are independent of every build choice.

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from conftest import annulus_points, make_build, sphere_points  # noqa: E402
from oracle.batched_apply import BatchedApplyCPU  # noqa: E402
from oracle.direct import direct_sum_c  # noqa: E402
from skelfmm_amd import operators as ops  # noqa: E402


def case(fname, X, op, kappa, tol, leaf, normals=None, seed=1):
    rs = np.random.RandomState(seed)
    q = rs.randn(X.shape[0])
    if op.is_complex:
        q = q + 1j * rs.randn(X.shape[0])
    gt = normals if op.geometry_side == "target" else None
    gs = normals if op.geometry_side == "source" else None
    yd = direct_sum_c(op.name, X, X, q, kappa, geom_trg=gt, geom_src=gs)
    fmm = make_build(X, op, kappa=kappa, tol=tol, leaf_size=leaf, geometry=normals)
    ya = BatchedApplyCPU(fmm, bytes_budget=1e8).matvec(q).numpy()
    extra = {} if normals is None else {"normals": normals}
    np.savez_compressed(os.path.join(HERE, fname), operator=op.name, kappa=kappa, tol=tol, leaf_size=leaf,
                        points=X, charges=q, direct=yd, apply=ya, **extra)
    err = np.linalg.norm(ya - yd) / np.linalg.norm(yd)
    print(f"{fname}: N={X.shape[0]} rel-l2(apply, direct) = {err:.2e}")


if __name__ == "__main__":
    case("laplace_2d_square.npz", np.random.default_rng(0).random((1500, 2)), ops.LAPLACE_2D, 0.0, 1e-6, 40)
    case("laplace_2d_annulus.npz", annulus_points(1500, 4), ops.LAPLACE_2D, 0.0, 1e-5, 40)
    case("laplace_3d_cube.npz", np.random.default_rng(0).random((2000, 3)), ops.LAPLACE_3D, 0.0, 1e-4, 100)
    case("helmholtz_3d_cube.npz", np.random.default_rng(1).random((1200, 3)), ops.HELMHOLTZ_3D, 4.0, 1e-4, 100)
    v = sphere_points(1200, 5)
    case("helmholtz_dl_sphere.npz", v, ops.HELMHOLTZ_DOUBLE_LAYER_3D, 3.0, 1e-4, 100, normals=v)
    case("laplace_adl_sphere.npz", v, ops.LAPLACE_ADJOINT_DOUBLE_LAYER_3D, 0.0, 1e-4, 100, normals=v)
