"""GPU: Helmholtz on the unit sphere at kappa = 10 pi, tol 1e-5, leaf 256 (the surface workload of
the benchmark, 1M points there, 20k-40k here so the oracles finish in seconds).

* complex128: the HIP apply against the oracle's restatement of ``SkelFMMBatched.matvec``
  (``skelfmm/batched_fmm.py:539-564``) on the same build, <= 1e-12 relative l2;
* complex64 (precision form): relative l2 against the FP64 direct sum (C oracle) at 1,000
  targets drawn by ``default_rng(0)``, with both sine/cosine paths of the FP32 kernels (the
  branch-free reduction when the host bounds kappa r, the library's sincosf otherwise).
  Bar: complex128 within tol; complex64 within 2 tol (FP32 rounding of kappa r alone costs
  ~kappa * diameter * 6e-8 ~ 4e-6 of phase on the far translations at this kappa).
* double layers with normals = positions (``skelfmm/util_batched.py:458-484,537-563``): on the
  surface rn = (x - y) . n ~ r^2 / 2, so the precision form forms rn in FP64 (otherwise the
  FP32 coordinates' rounding, ~6e-8 of a box side, costs ~5e-4 at 1M points). Checked against
  the oracle's restatement of the reference's single-precision form on the same build.
"""

import numpy as np
import pytest
import torch

from conftest import make_build, rel_l2, sphere_points
from oracle.batched_apply import BatchedApplyCPU
from oracle.direct import direct_sum_c
from skelfmm_amd import operators as ops

pytestmark = pytest.mark.gpu

KAPPA = 10 * np.pi
TOL = 1e-5
OPS = {"helmholtz_3d": ops.HELMHOLTZ_3D, "helmholtz_double_layer_3d": ops.HELMHOLTZ_DOUBLE_LAYER_3D,
       "helmholtz_adjoint_double_layer_3d": ops.HELMHOLTZ_ADJOINT_DOUBLE_LAYER_3D}


def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    return torch.device("cuda", 0)


def _charges(n):
    rs = np.random.RandomState(1)
    re = rs.randn(n)
    return re + 1j * rs.randn(n)


def _build(name, n, tol=TOL, leaf=256):
    X = sphere_points(n, 0)
    op = OPS[name]
    return X, make_build(X, op, kappa=KAPPA, tol=tol, leaf_size=leaf, geometry=X if op.geometry_side else None)


def _apply(fmm, q, single):
    from skelfmm_amd.batched import SkelFMMBatched

    dev = SkelFMMBatched(fmm.tree, fmm.operator, kappa=fmm.kappa).build_from_copy(fmm, use_single_precision=single)
    x = torch.as_tensor(q.astype(np.complex64) if single else q).to(_gpu())
    y = dev.matvec(x).cpu().numpy().astype(np.complex128)
    return y, dev, x


@pytest.mark.parametrize("name", ["helmholtz_3d", "helmholtz_double_layer_3d"])
def test_sphere_complex128_matches_oracle(name):
    X, fmm = _build(name, 20000)
    print(f"{name} sphere N={X.shape[0]} kappa=10pi levels={fmm.nlevels} max rank={int(fmm.rank_list.max())}",
          flush=True)
    q = _charges(X.shape[0])
    y, _, _ = _apply(fmm, q, single=False)
    ref = BatchedApplyCPU(fmm, bytes_budget=2e8).matvec(q).numpy()
    err = rel_l2(y, ref)
    print(f"  complex128 GPU vs oracle: {err:.2e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("name", ["helmholtz_3d", "helmholtz_double_layer_3d", "helmholtz_adjoint_double_layer_3d"])
def test_sphere_complex64_vs_direct_sum(name):
    X, fmm = _build(name, 40000)
    n = X.shape[0]
    q = _charges(n)
    idx = np.random.default_rng(0).choice(n, 1000, replace=False)
    side = OPS[name].geometry_side
    exact = direct_sum_c(name, X[idx], X, q, kappa=KAPPA, geom_trg=X[idx] if side == "target" else None,
                         geom_src=X if side == "source" else None)
    y128, _, _ = _apply(fmm, q, single=False)
    y64, d32, x32 = _apply(fmm, q, single=True)
    assert 0 < KAPPA * d32.max_distance <= 1600  # the branch-free sine/cosine path ran
    d32.max_distance = 0.0  # library sincosf path
    d32._invalidate()
    y64_lib = d32.matvec(x32).cpu().numpy().astype(np.complex128)
    e128, e64, e64_lib = (rel_l2(y[idx], exact) for y in (y128, y64, y64_lib))
    print(f"{name} sphere N={n}: vs FP64 direct sum at 1000 targets: complex128 {e128:.2e}, complex64 "
          f"{e64:.2e} (reduced sincos), {e64_lib:.2e} (sincosf)", flush=True)
    assert e128 <= TOL
    assert e64 <= 2 * TOL and e64_lib <= 2 * TOL
    assert rel_l2(y64, y64_lib) <= 1e-5


@pytest.mark.parametrize("name", ["helmholtz_double_layer_3d", "helmholtz_adjoint_double_layer_3d"])
def test_sphere_double_layer_precision_form_beats_reference_single(name):
    """Normals = positions: the precision form (rn in FP64) against the oracle's restatement of
    the reference's single-precision form (everything FP32, absolute coordinates), both measured
    against the complex128 apply of the same build."""
    X, fmm = _build(name, 8000, tol=1e-6, leaf=128)
    q = _charges(X.shape[0])
    y128, _, _ = _apply(fmm, q, single=False)
    y64, _, _ = _apply(fmm, q, single=True)
    ref32 = BatchedApplyCPU(fmm, use_single_precision=True, bytes_budget=2e8).matvec(
        q.astype(np.complex64)).numpy().astype(np.complex128)
    e_gpu, e_ref = rel_l2(y64, y128), rel_l2(ref32, y128)
    print(f"{name} sphere N={X.shape[0]} FP32: reference single-precision form {e_ref:.2e}, "
          f"GPU precision form {e_gpu:.2e}", flush=True)
    assert e_gpu <= 0.05 * e_ref
    assert e_gpu <= 2e-5
