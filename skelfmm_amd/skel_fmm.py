"""Public wrapper: drop-in for ``skelfmm.SkelFMM`` (reference ``skelfmm/skel_fmm.py:55-282``).

Contract kept from the reference: the constructor keywords and their defaults, the public
attributes (``kernel_backend``, ``device``, ``rank_or_tol``, ``leaf_size``, the
``*_time_seconds`` and ``*_storage_bytes`` figures, ``nbytes``, ``nlevels``, ``nboxes``,
``nleaves``, ``rank_max``), ``build()`` idempotent and returning ``self``, ``apply`` /
``matvec`` on a 1D host vector returning a host array of the stored precision, the
lazy-build notice, the error types and messages, ``relerr_fmm_apply_check`` (relative
l-infinity error on a sampled target subset) and ``benchmark_matvec`` (wall-clock timing
dict).

How it is organised here: a build is three timed phases (tree, skeletons, device layout)
run by ``_BuildPhases``; everything the build reports about itself is derived afterwards
from the objects it produced (``_report_storage``). The apply itself is the device plan of
``SkelFMMBatched`` (HIP kernels); this class only converts host arrays.

Opt-in extensions (keyword-only): ``multi_rhs=True`` lets ``apply`` take ``(N, R)``
right-hand sides; ``build_device="gpu"`` builds the tree, its lists and the work tables on the GPU and
runs the interpolative decompositions as batched HIP kernels; ``rank_model`` (a ``learned_rank.RankMLP`` or ``"default"``) switches those to the
learned fixed-rank sketched form with adaptive fallback; ``save(path)`` /
``SkelFMM.load(path)`` write and read the built operator (``skelfmm_amd.serialize``).
"""

from __future__ import annotations

import contextlib
import io
import time

import numpy as np
import torch

from .build import RecursiveSkeletonization
from .tree import BalancedTree

__all__ = ["SkelFMM", "warmup_cuda_fmm_apply"]

# (input is complex, stored precision is FP32) -> torch dtype of the device vectors
_VECTOR_DTYPES = {
    (False, False): torch.float64,
    (False, True): torch.float32,
    (True, False): torch.complex128,
    (True, True): torch.complex64,
}


def _batched_tensor_dtype(x, real_dtype):
    """Device dtype for host array ``x`` when the build stores ``real_dtype`` (reference
    ``skel_fmm.py:32-35``): real stays real, complex becomes the complex type of that width."""
    return _VECTOR_DTYPES[(bool(np.iscomplexobj(x)), real_dtype == torch.float32)]


def _on_gpu(dev):
    return isinstance(dev, torch.device) and dev.type == "cuda"


def warmup_cuda_fmm_apply(apply, x):
    """Run one untimed apply of ``apply`` (anything with ``.apply`` and ``.device``) and wait
    for the device, so that later timings exclude first-call costs (reference
    ``skel_fmm.py:38-52``). Returns False, doing nothing, when the operator is not on a GPU."""
    dev = getattr(apply, "device", None)
    if not _on_gpu(dev):
        return False
    apply.apply(x)
    waiter = getattr(apply, "_synchronize", None) or (lambda: torch.cuda.synchronize(dev))
    waiter()
    return True


class _BuildPhases:
    """Wall-clock seconds of named build phases plus the whole."""

    def __init__(self):
        self.seconds = {}
        self._t_start = None

    @contextlib.contextmanager
    def phase(self, name):
        t = time.perf_counter()
        try:
            yield
        finally:
            self.seconds[name] = time.perf_counter() - t

    def __enter__(self):
        self._t_start = time.perf_counter()
        return self

    def __exit__(self, *exc):
        self.seconds["total"] = time.perf_counter() - self._t_start
        return False


def _muted(quiet):
    """Swallow stdout/stderr of the build when ``quiet`` (the reference's ``quiet=True``)."""
    if not quiet:
        return contextlib.nullcontext()
    stack = contextlib.ExitStack()
    stack.enter_context(contextlib.redirect_stdout(io.StringIO()))
    stack.enter_context(contextlib.redirect_stderr(io.StringIO()))
    return stack


def _host_build_bytes(fmm):
    """Bytes the host build holds: candidate ids, interpolation matrices, box tables."""
    if fmm is None:
        return 0
    arrays = list(fmm.cand) + [fmm.bs_list, fmm.rank_list, fmm.tree_boxes, fmm.tree_lev_sep]
    return int(fmm.nbytes_proj + sum(a.nbytes for a in arrays))


def _largest_rank(fmm):
    if fmm is None:
        return 0
    ranks = np.asarray(fmm.rank_list)
    return int(ranks.max(initial=0))


class SkelFMM:
    """A built skeletonized-FMM operator whose apply runs on the GPU through HIP kernels."""

    _TIMES = ("tree_time_seconds", "skel_time_seconds", "setup_time_seconds", "build_time_seconds")
    _SIZES = ("host_build_storage_bytes", "host_storage_bytes", "device_storage_bytes", "total_storage_bytes",
              "nbytes")

    def __init__(self, points, operator, geometry=None, tol=1e-5, rank_or_tol=None, leaf_size=200,
                 use_single_precision=False, param=0.0, quiet=True, keep_host_fmm=False, defer_build=False,
                 npoints_max=500, p="auto", max_bs=-1, max_rank=-1, *, multi_rhs=False, build_device="cpu",
                 rank_model=None):
        if build_device not in ("cpu", "gpu"):
            raise ValueError("build_device must be 'cpu' or 'gpu'")
        # what the caller asked for
        self.points = points
        self.operator = operator
        self.geometry = geometry
        self.param = param
        self.rank_or_tol = rank_or_tol if rank_or_tol is not None else tol
        self.leaf_size = leaf_size
        self.build_device = build_device
        self.rank_model = rank_model
        self._quiet = quiet
        self._keep_host_fmm = keep_host_fmm
        self._use_single_precision = use_single_precision
        self._multi_rhs = multi_rhs
        self._skel_options = {"npoints_max": npoints_max, "p": p}
        self._size_caps = {"max_bs": max_bs, "max_rank": max_rank}
        # what the build produces
        self.tree = None
        self.fmm = None
        self.batched = None
        self.device = None
        self.kernel_backend = None
        self.nlevels = self.nboxes = self.nleaves = 0
        self.rank_max = 0
        self.build_stats = {}
        for name in self._TIMES:
            setattr(self, name, 0.0)
        for name in self._SIZES:
            setattr(self, name, 0)
        self._is_built = False
        self._lazy_build_notice_printed = False
        if not defer_build:
            self.build()

    # ------------------------------------------------------------------ build
    def _resolve_rank_model(self):
        model = self.rank_model
        if isinstance(model, str) and model == "default":
            from .learned_rank import load_default

            model = load_default()
            if model is None:
                raise RuntimeError("rank_model='default' but no trained rank network is installed")
        if model is not None and self.build_device != "gpu":
            raise ValueError("rank_model requires build_device='gpu'")
        return model

    def _make_tree(self):
        if self.build_device == "gpu":  # the same tree (bit for bit), built by HIP kernels
            from .gpu_tree import GPUBalancedTree

            self.tree = GPUBalancedTree(self.points, leaf_size=self.leaf_size)
        else:
            self.tree = BalancedTree(self.points, leaf_size=self.leaf_size)
        self.nlevels, self.nboxes, self.nleaves = self.tree.nlevels, self.tree.nboxes, self.tree.nleaves

    def _make_skeletons(self, npoints_max, p, verbose):
        self.fmm = RecursiveSkeletonization(self.tree, self.operator, kappa=self.param, tol=self.rank_or_tol,
                                            geometry=self.geometry, **self._size_caps)
        self.fmm.skel_tree(npoints_max=npoints_max, p=p, verbose=verbose, device=self.build_device,
                           rank_model=self._resolve_rank_model())
        if self.build_device == "gpu":
            torch.cuda.synchronize()  # the phase ends when the device IDs are done
        self.build_stats = dict(self.fmm.gpu_stats)

    def _make_device_layout(self, verbose):
        from .batched import SkelFMMBatched

        self.fmm.setup_lists(verbose=verbose)
        self.batched = SkelFMMBatched(self.tree, self.operator, kappa=self.param).build_from_copy(
            self.fmm, use_single_precision=self._use_single_precision)

    def _report_storage(self):
        host_build = _host_build_bytes(self.fmm)
        self.rank_max = _largest_rank(self.fmm)
        if not self._keep_host_fmm:
            self.fmm = None
        self.host_build_storage_bytes = host_build
        self.host_storage_bytes = host_build if self.fmm is not None else 0
        self.device_storage_bytes = int(self.batched.nbytes_proj() + self.batched.nbytes_lists())
        self.total_storage_bytes = self.host_storage_bytes + self.device_storage_bytes
        self.nbytes = self.total_storage_bytes

    def build(self, *, npoints_max=None, p=None, verbose=False):
        """Tree, interpolative decompositions, device layout. A second call does nothing."""
        if self._is_built:
            return self
        opts = dict(self._skel_options)
        if npoints_max is not None:
            opts["npoints_max"] = npoints_max
        if p is not None:
            opts["p"] = p
        with _BuildPhases() as phases, _muted(self._quiet):
            with phases.phase("tree"):
                self._make_tree()
            with phases.phase("skel"):
                self._make_skeletons(opts["npoints_max"], opts["p"], verbose)
            with phases.phase("setup"):
                self._make_device_layout(verbose)
        for name in ("tree", "skel", "setup"):
            setattr(self, f"{name}_time_seconds", phases.seconds[name])
        self.build_time_seconds = phases.seconds["total"]
        self.device = self.batched.device
        self.kernel_backend = self.batched.kernel_backend
        self._report_storage()
        self._is_built = True
        return self

    # --------------------------------------------------------- serialization
    def save(self, path):
        """Write the built operator to ``path`` (versioned ``.npz``; ``skelfmm_amd.serialize``)."""
        self._ensure_built()
        meta = {"nboxes": int(self.nboxes), "nleaves": int(self.nleaves), "rank_max": int(self.rank_max),
                "build_device": self.build_device, "build_time_seconds": float(self.build_time_seconds)}
        self.batched.save(path, points=np.asarray(self.points), geometry=self.geometry, leaf_size=self.leaf_size,
                          extra=meta)

    @classmethod
    def load(cls, path, *, use_single_precision=False, multi_rhs=False, quiet=True, verify=True):
        """A built ``SkelFMM`` from a file written by ``save``: no tree, IDs or lists are formed."""
        from . import operators as ops_mod
        from .batched import SkelFMMBatched
        from .serialize import load_layout

        t_start = time.perf_counter()
        rec = load_layout(path, verify=verify)
        head = rec.header
        known = {op.name: op for op in ops_mod.BUILTIN_OPERATORS}
        if head["operator"] not in known:
            raise ValueError(f"{path}: unknown operator {head['operator']!r}")
        obj = cls(rec.points, known[head["operator"]], geometry=rec.geometry, tol=head["tol"],
                  leaf_size=head["leaf_size"], use_single_precision=use_single_precision, param=head["kappa"],
                  quiet=quiet, defer_build=True, multi_rhs=multi_rhs)
        meta = head.get("extra", {})
        obj.batched = SkelFMMBatched.from_layout(rec.layout, obj.operator, kappa=head["kappa"],
                                                 use_single_precision=use_single_precision, tol=head["tol"],
                                                 nlevels=meta.get("nlevels"))
        obj.nlevels = meta.get("nlevels") or 0
        obj.nboxes = meta.get("nboxes", int(rec.layout.nboxes))
        obj.nleaves = meta.get("nleaves", 0)
        obj.rank_max = meta.get("rank_max", 0)
        obj.build_device = meta.get("build_device", "cpu")
        obj.device = obj.batched.device
        obj.kernel_backend = obj.batched.kernel_backend
        obj.device_storage_bytes = obj.total_storage_bytes = obj.nbytes = int(
            obj.batched.nbytes_proj() + obj.batched.nbytes_lists())
        obj.setup_time_seconds = obj.build_time_seconds = time.perf_counter() - t_start
        obj._is_built = True
        return obj

    # ------------------------------------------------------------------ apply
    def _synchronize(self):
        if _on_gpu(self.device):
            torch.cuda.synchronize(self.device)

    def _ensure_built(self, *, announce=False):
        if getattr(self, "_is_built", True):
            return
        if announce and not self._lazy_build_notice_printed:
            self._lazy_build_notice_printed = True
            print("\tBuilding SkelFMM on first matvec; subsequent matvecs will be fast.")
        self.build()

    def apply(self, x):
        """``y = K~ x`` for a host vector ``x`` of length N; returns a host array of the stored
        precision (real stays real, complex becomes complex64 / complex128)."""
        self._ensure_built(announce=True)
        host = np.asarray(x)
        shape_ok = host.ndim == 1 or (self._multi_rhs and host.ndim == 2)
        if not shape_ok:
            raise ValueError("apply expects a 1D vector")
        dtype = _batched_tensor_dtype(host, self.batched.real_dtype)
        on_device = torch.as_tensor(host).to(device=self.device, dtype=dtype)
        return self.batched.matvec(on_device).cpu().numpy()

    def matvec(self, x):
        return self.apply(x)

    # ---------------------------------------------------------- diagnostics
    def _dense_rows(self, rows):
        """Rows ``rows`` of the exact kernel matrix, geometry on the operator's side."""
        side = self.operator.geometry_side
        g_trg = self.geometry[rows] if side == "target" else None
        g_src = self.geometry if side == "source" else None
        return self.operator.numpy_matrix(self.points[rows], self.points, geom_trg=g_trg, geom_src=g_src,
                                          param=self.param)

    def relerr_fmm_apply_check(self, x, *, ncheck=64, seed=0, indices=None):
        """Relative l-infinity error of the fast apply on a target subset,
        ``max|y_dense - y_fmm| / max|y_dense|`` (reference ``skel_fmm.py:218-257``). Without
        ``indices``, ``min(ncheck, N)`` distinct targets are drawn by ``default_rng(seed)``."""
        self._ensure_built()
        x = np.asarray(x)
        if x.ndim != 1:
            raise ValueError("relerr_fmm_apply_check expects a 1D source vector")
        npts = self.points.shape[0]
        if x.shape[0] != npts:
            raise ValueError("Source vector length must match the number of points")
        if indices is None:
            rows = np.random.default_rng(seed).choice(npts, size=min(int(ncheck), npts), replace=False)
        else:
            rows = np.asarray(indices, dtype=int)
        fast = self.apply(x)[rows]
        exact = self._dense_rows(rows) @ x
        scale = max(float(np.abs(exact).max(initial=0.0)), 1e-30)
        return float(np.abs(exact - fast).max(initial=0.0)) / scale

    def benchmark_matvec(self, *, nrepeat=3, warmup=1, seed=0):
        """Wall-clock seconds of ``apply`` with host numpy in and out on a ``RandomState(seed)``
        normal vector, after a device warm-up (reference ``skel_fmm.py:259-282``)."""
        self._ensure_built()
        x = np.random.RandomState(seed).randn(self.points.shape[0])
        warmup_cuda_fmm_apply(self, x)

        def timed():
            self._synchronize()
            t = time.perf_counter()
            self.apply(x)
            self._synchronize()
            return time.perf_counter() - t

        for _ in range(max(warmup, 0)):
            timed()
        samples = np.array([timed() for _ in range(max(nrepeat, 1))])
        return {"times_seconds": tuple(float(t) for t in samples), "avg_seconds": float(samples.mean()),
                "min_seconds": float(samples.min()), "max_seconds": float(samples.max())}
