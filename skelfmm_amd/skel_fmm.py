"""Public wrapper (drop-in for ``skelfmm.SkelFMM``, ``skelfmm/skel_fmm.py:55-282``).

``SkelFMM(points, operator, tol, leaf_size).build()`` builds the tree and the
skeletons on the host and uploads the box layout to the GPU; ``apply(x)``
runs the HIP apply and returns a host numpy array. Constructor keywords,
attributes, error types, the lazy-build notice, ``relerr_fmm_apply_check``
and ``benchmark_matvec`` follow the reference. Opt-in extension: ``apply``
also accepts an ``(N, R)`` array of right-hand sides when ``multi_rhs=True``;
``build_device="gpu"`` runs the per-box interpolative decompositions as batched
HIP kernels, and ``rank_model`` (a ``learned_rank.RankMLP`` or ``"default"``)
switches them to the learned fixed-rank sketched form with adaptive fallback;
``save(path)`` / ``SkelFMM.load(path)`` write and read the built operator as one
versioned file (``skelfmm_amd.serialize``).
"""

from __future__ import annotations

import contextlib
import io
from time import perf_counter

import numpy as np
import torch

from .build import RecursiveSkeletonization
from .tree import BalancedTree

__all__ = ["SkelFMM", "warmup_cuda_fmm_apply"]


def _batched_tensor_dtype(x, real_dtype):
    """Apply dtype of a host array (``skel_fmm.py:32-35``)."""
    if np.iscomplexobj(x):
        return torch.complex64 if real_dtype == torch.float32 else torch.complex128
    return real_dtype


def warmup_cuda_fmm_apply(apply, x):
    """One untimed apply plus a device sync (``skel_fmm.py:38-52``)."""
    device = getattr(apply, "device", None)
    if not (isinstance(device, torch.device) and device.type == "cuda"):
        return False
    apply.apply(x)
    sync = getattr(apply, "_synchronize", None)
    if sync is not None:
        sync()
    else:
        torch.cuda.synchronize(device)
    return True


class SkelFMM:
    """Build and apply a fast operator through the batched HIP interface."""

    def __init__(self, points, operator, geometry=None, tol=1e-5, rank_or_tol=None, leaf_size=200,
                 use_single_precision=False, param=0.0, quiet=True, keep_host_fmm=False, defer_build=False,
                 npoints_max=500, p="auto", max_bs=-1, max_rank=-1, multi_rhs=False, build_device="cpu",
                 rank_model=None):
        self.points = points
        self.operator = operator
        self.geometry = geometry
        self.param = param
        self.rank_or_tol = tol if rank_or_tol is None else rank_or_tol
        self.leaf_size = leaf_size
        self.tree = None
        self.nlevels = 0
        self.nboxes = 0
        self.nleaves = 0
        self._quiet = quiet
        self._keep_host_fmm = keep_host_fmm
        self._use_single_precision = use_single_precision
        self._build_npoints_max = npoints_max
        self._build_p = p
        self._max_bs_limit = max_bs
        self._max_rank_limit = max_rank
        self._multi_rhs = multi_rhs
        if build_device not in ("cpu", "gpu"):
            raise ValueError("build_device must be 'cpu' or 'gpu'")
        self.build_device = build_device
        self.rank_model = rank_model
        self.build_stats = {}
        self._is_built = False
        self._lazy_build_notice_printed = False

        self.fmm = None
        self.batched = None
        self.device = None
        self.kernel_backend = None
        self.tree_time_seconds = 0.0
        self.skel_time_seconds = 0.0
        self.setup_time_seconds = 0.0
        self.build_time_seconds = 0.0
        self.rank_max = 0
        self.host_build_storage_bytes = 0
        self.host_storage_bytes = 0
        self.device_storage_bytes = 0
        self.total_storage_bytes = 0
        self.nbytes = 0

        if not defer_build:
            self.build()

    def _estimate_host_storage_bytes(self):
        if self.fmm is None:
            return 0
        f = self.fmm
        total = sum(c.nbytes for c in f.cand) + f.nbytes_proj + f.bs_list.nbytes + f.rank_list.nbytes
        total += f.tree_boxes.nbytes + f.tree_lev_sep.nbytes
        return int(total)

    def _compute_rank_max(self):
        if self.fmm is None:
            return 0
        r = np.asarray(self.fmm.rank_list, dtype=int)
        r = r[r > 0]
        return int(r.max()) if r.size else 0

    @staticmethod
    def _synchronize_default():
        torch.cuda.synchronize()

    def _synchronize(self):
        if isinstance(self.device, torch.device) and self.device.type == "cuda":
            torch.cuda.synchronize(self.device)

    def build(self, *, npoints_max=None, p=None, verbose=False):
        if self._is_built:
            return self
        if npoints_max is None:
            npoints_max = self._build_npoints_max
        if p is None:
            p = self._build_p

        build_t0 = perf_counter()
        ctx = contextlib.ExitStack()
        with ctx as stack:
            if self._quiet:
                stack.enter_context(contextlib.redirect_stdout(io.StringIO()))
                stack.enter_context(contextlib.redirect_stderr(io.StringIO()))
            t0 = perf_counter()
            self.tree = BalancedTree(self.points, leaf_size=self.leaf_size)
            self.tree_time_seconds = perf_counter() - t0
            self.nlevels = self.tree.nlevels
            self.nboxes = self.tree.nboxes
            self.nleaves = self.tree.nleaves
            self.fmm = RecursiveSkeletonization(self.tree, self.operator, kappa=self.param, tol=self.rank_or_tol,
                                                max_bs=self._max_bs_limit, max_rank=self._max_rank_limit,
                                                geometry=self.geometry)
            model = self.rank_model
            if isinstance(model, str) and model == "default":
                from .learned_rank import load_default

                model = load_default()
                if model is None:
                    raise RuntimeError("rank_model='default' but no trained rank network is installed")
            if model is not None and self.build_device != "gpu":
                raise ValueError("rank_model requires build_device='gpu'")
            t0 = perf_counter()
            self.fmm.skel_tree(npoints_max=npoints_max, p=p, verbose=verbose, device=self.build_device,
                               rank_model=model)
            if self.build_device == "gpu":
                self._synchronize_default()
            self.skel_time_seconds = perf_counter() - t0
            self.build_stats = dict(self.fmm.gpu_stats)

            t0 = perf_counter()
            self.fmm.setup_lists(verbose=verbose)
            from .batched import SkelFMMBatched

            self.batched = SkelFMMBatched(self.tree, self.operator, kappa=self.param)
            self.batched.build_from_copy(self.fmm, use_single_precision=self._use_single_precision)
            self.setup_time_seconds = perf_counter() - t0
        self.build_time_seconds = perf_counter() - build_t0

        self.device = self.batched.device
        self.kernel_backend = self.batched.kernel_backend
        host_build = self._estimate_host_storage_bytes()
        self.rank_max = self._compute_rank_max()
        if self._keep_host_fmm:
            host_storage = host_build
        else:
            self.fmm = None
            host_storage = 0
        device_storage = int(self.batched.nbytes_proj() + self.batched.nbytes_lists())
        self.host_build_storage_bytes = host_build
        self.host_storage_bytes = host_storage
        self.device_storage_bytes = device_storage
        self.total_storage_bytes = host_storage + device_storage
        self.nbytes = self.total_storage_bytes
        self._is_built = True
        return self

    # --------------------------------------------------------- serialization
    def save(self, path):
        """Write the built operator to ``path`` (versioned ``.npz``; ``skelfmm_amd.serialize``)."""
        self._ensure_built()
        self.batched.save(path, points=np.asarray(self.points), geometry=self.geometry, leaf_size=self.leaf_size,
                          extra={"nboxes": int(self.nboxes), "nleaves": int(self.nleaves),
                                 "rank_max": int(self.rank_max), "build_device": self.build_device,
                                 "build_time_seconds": float(self.build_time_seconds)})

    @classmethod
    def load(cls, path, *, use_single_precision=False, multi_rhs=False, quiet=True, verify=True):
        """A built ``SkelFMM`` from a file written by ``save``: no tree, IDs or lists are formed."""
        from . import operators as ops_mod
        from .batched import SkelFMMBatched
        from .serialize import load_layout

        t0 = perf_counter()
        rec = load_layout(path, verify=verify)
        h = rec.header
        by_name = {op.name: op for op in ops_mod.BUILTIN_OPERATORS}
        if h["operator"] not in by_name:
            raise ValueError(f"{path}: unknown operator {h['operator']!r}")
        self = cls(rec.points, by_name[h["operator"]], geometry=rec.geometry, tol=h["tol"],
                   leaf_size=h["leaf_size"], use_single_precision=use_single_precision, param=h["kappa"],
                   quiet=quiet, defer_build=True, multi_rhs=multi_rhs)
        ex = h.get("extra", {})
        self.batched = SkelFMMBatched.from_layout(rec.layout, self.operator, kappa=h["kappa"],
                                                  use_single_precision=use_single_precision, tol=h["tol"],
                                                  nlevels=ex.get("nlevels"))
        self.nlevels = ex.get("nlevels") or 0
        self.nboxes = ex.get("nboxes", int(rec.layout.nboxes))
        self.nleaves = ex.get("nleaves", 0)
        self.rank_max = ex.get("rank_max", 0)
        self.build_device = ex.get("build_device", "cpu")
        self.device = self.batched.device
        self.kernel_backend = self.batched.kernel_backend
        self.device_storage_bytes = int(self.batched.nbytes_proj() + self.batched.nbytes_lists())
        self.total_storage_bytes = self.device_storage_bytes
        self.nbytes = self.total_storage_bytes
        self.setup_time_seconds = perf_counter() - t0
        self.build_time_seconds = self.setup_time_seconds
        self._is_built = True
        return self

    def _ensure_built(self, *, announce=False):
        if getattr(self, "_is_built", True):
            return
        if announce and not self._lazy_build_notice_printed:
            print("\tBuilding SkelFMM on first matvec; subsequent matvecs will be fast.")
            self._lazy_build_notice_printed = True
        self.build()

    def apply(self, x):
        self._ensure_built(announce=True)
        x_np = np.asarray(x)
        if x_np.ndim != 1 and not (self._multi_rhs and x_np.ndim == 2):
            raise ValueError("apply expects a 1D vector")
        x_t = torch.as_tensor(x_np, dtype=_batched_tensor_dtype(x_np, self.batched.real_dtype))
        y_t = self.batched.matvec(x_t.to(self.device))
        return y_t.detach().cpu().numpy()

    def matvec(self, x):
        return self.apply(x)

    def relerr_fmm_apply_check(self, x, *, ncheck=64, seed=0, indices=None):
        """Relative l_inf subset error ``||y_dense - y_fmm||_inf / ||y_dense||_inf`` (``skel_fmm.py:218-257``)."""
        self._ensure_built()
        x = np.asarray(x)
        if x.ndim != 1:
            raise ValueError("relerr_fmm_apply_check expects a 1D source vector")
        if x.shape[0] != self.points.shape[0]:
            raise ValueError("Source vector length must match the number of points")
        if indices is None:
            rng = np.random.default_rng(seed)
            indices = rng.choice(self.points.shape[0], size=min(int(ncheck), self.points.shape[0]), replace=False)
        else:
            indices = np.asarray(indices, dtype=int)
        geom_src = geom_trg = None
        if self.operator.geometry_side == "source":
            geom_src = self.geometry
        elif self.operator.geometry_side == "target":
            geom_trg = self.geometry[indices]
        y_fast = self.apply(x)
        y_true = self.operator.numpy_matrix(self.points[indices], self.points, geom_trg=geom_trg, geom_src=geom_src,
                                            param=self.param) @ x
        denom = max(np.linalg.norm(y_true, ord=np.inf), 1e-30)
        return np.linalg.norm(y_true - y_fast[indices], ord=np.inf) / denom

    def benchmark_matvec(self, *, nrepeat=3, warmup=1, seed=0):
        """Wall-clock apply timing, host numpy in and out (``skel_fmm.py:259-282``)."""
        self._ensure_built()
        x = np.random.RandomState(seed).randn(self.points.shape[0])
        warmup_cuda_fmm_apply(self, x)
        times = []
        for _ in range(max(warmup, 0)):
            self.apply(x)
            self._synchronize()
        for _ in range(max(nrepeat, 1)):
            self._synchronize()
            t0 = perf_counter()
            self.apply(x)
            self._synchronize()
            times.append(perf_counter() - t0)
        return {
            "times_seconds": tuple(times),
            "avg_seconds": float(np.mean(times)),
            "min_seconds": float(np.min(times)),
            "max_seconds": float(np.max(times)),
        }
