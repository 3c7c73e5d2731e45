"""Device-resident apply (drop-in for ``skelfmm.batched_fmm.SkelFMMBatched``).

``build_from_copy`` flattens a host build into the ragged box layout
(``skelfmm_amd.layout``) and uploads it once; ``matvec`` runs the whole apply
as native HIP launches through ``skf_apply`` (charge gather, upward
interpolation per level, one near-field launch for all lists, downward
interpolation per level, result scatter) on PyTorch's current stream. Nothing
of the apply runs on the CPU. Reference: ``skelfmm/batched_fmm.py:6-564``.

Forms:

* plain (``use_single_precision=False``): FP64 / complex128 throughout;
* precision (``use_single_precision=True``): FP32 / complex64 points relative to
  their box centre (centres kept in FP64), FP32 interpolation matrices and
  vectors; near-field sums are formed per 64-source tile in FP32 and
  accumulated in FP64 across tiles.

Extension over the reference (opt-in): ``matvec`` accepts an ``(N, R)`` block
of right-hand sides.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import hip, operators as operators_mod
from .layout import ITEM_ORDERS, build_layout, merge_layouts, order_items

__all__ = ["SkelFMMBatched", "BatchedInstances", "CapturedApply"]


class SkelFMMBatched(torch.nn.Module):
    def __init__(self, tree, operator, kappa: float = 0.0):
        super().__init__()
        self.N = tree.N
        self.tree = tree
        self.nlevels = tree.nlevels
        self.nboxes = tree.nboxes
        self.operator = operators_mod._require_operator(operator)
        self.kappa = float(kappa)
        self.ndim = tree.ndim
        self.level_sep = tree.level_sep
        self.kernel_backend = "hip"
        if self.operator.kernel_id < 0:
            raise NotImplementedError(
                "the HIP apply supports the built-in operators of skelfmm_amd.operators only")
        self._plans = {}
        self._graphs = {}
        self.generation = 0

    # ------------------------------------------------------------------ build
    @classmethod
    def from_layout(cls, L, operator, kappa=0.0, use_single_precision=False, device=None, tol=None, nlevels=None):
        """Device apply from a box layout alone, e.g. one read back by ``serialize.load_layout``."""
        self = cls.__new__(cls)
        torch.nn.Module.__init__(self)
        self.operator = operators_mod._require_operator(operator)
        if self.operator.kernel_id < 0:
            raise NotImplementedError("the HIP apply supports the built-in operators of skelfmm_amd.operators only")
        self.kappa = float(kappa)
        self.ndim = int(L.ndim)
        self.N = int(L.N)
        self.nboxes = int(L.nboxes)
        self.nlevels = nlevels
        self.tree = None
        self.tol = tol
        self.root_level = int(L.root_level)
        self.kernel_backend = "hip"
        self._plans = {}
        self._graphs = {}
        self.generation = 0
        return self.upload_layout(L, use_single_precision, device)

    def save(self, path, *, points=None, geometry=None, leaf_size=-1, extra=None, extra_arrays=None):
        """Write this apply's layout and operator to a versioned ``.npz`` (``serialize.save_layout``).

        ``points`` (caller's order) default to the layout's own coordinates."""
        from .serialize import save_layout

        L = self.layout
        if points is None:
            points = np.empty((L.N, L.ndim))
            points[L.leaf_user] = L.points[:, L.leaf_slot].T
        if geometry is None and L.normals is not None:
            geometry = np.empty((L.N, 3))
            geometry[L.leaf_user] = L.normals[:, L.leaf_slot].T
        save_layout(path, L, operator_name=self.operator.name, kappa=self.kappa,
                    tol=-1.0 if getattr(self, "tol", None) is None else self.tol, leaf_size=leaf_size,
                    points=points, geometry=geometry, extra=dict(extra or {}, nlevels=self.nlevels),
                    extra_arrays=extra_arrays)
    def build_from_copy(self, skel_fmm, use_single_precision=False, device=None, max_sources=16384):
        self.tol = skel_fmm.tol
        self.root_level = skel_fmm.root_level
        self.max_rank = skel_fmm.max_rank
        self.max_bs = skel_fmm.max_bs
        self.min_rank = skel_fmm.min_rank
        self.max_rank_leaf = skel_fmm.max_rank_leaf
        self.max_bs_leaf = skel_fmm.max_bs_leaf
        self.rank_list = torch.as_tensor(skel_fmm.rank_list)
        self.bs_list = torch.as_tensor(skel_fmm.bs_list)
        nf_dev = None
        if getattr(skel_fmm, "build_device", "cpu") == "gpu":  # a GPU build also forms its work tables there
            hip.require_gpu()
            nf_dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        L = build_layout(skel_fmm, max_sources=max_sources, near_field_device=nf_dev)
        return self.upload_layout(L, use_single_precision, device)

    def upload_layout(self, L, use_single_precision=False, device=None):
        """Copy a box layout to the GPU once; applies then reuse it."""
        hip.require_gpu()
        hip.lib()
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        self.use_single_precision = bool(use_single_precision)
        self.precision = 1 if use_single_precision else 0
        self.real_dtype = torch.float32 if use_single_precision else torch.float64
        self.complex_dtype = torch.complex64 if use_single_precision else torch.complex128
        self.mfma_min_channels = 0
        self._level_variants = {}
        self.layout = L
        dev = self.device
        rd = self.real_dtype

        def i32(a):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)

        def i64(a):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int64), device=dev)

        if self.precision == 0:
            pts = torch.as_tensor(L.points, dtype=torch.float64, device=dev)
        else:
            rel = L.points - L.centers[L.slot_box, :self.ndim].T
            pts = torch.as_tensor(rel, dtype=torch.float32, device=dev)
        self.XX_slots = pts.contiguous()
        self.normals = None
        self.geo64 = None
        if L.normals is not None:
            self.normals = torch.as_tensor(L.normals, dtype=rd, device=dev).contiguous()
            if self.precision == 1:
                # double layers in the precision form: the normal projection (x - y) . n is formed in
                # FP64 from absolute coordinates, normals and n . x per slot (rows of 7 x nslots)
                P64 = np.zeros((3, L.nslots))
                P64[:self.ndim] = L.points
                N64 = np.asarray(L.normals, dtype=np.float64)
                geo = np.concatenate([P64, N64, (P64 * N64).sum(axis=0, keepdims=True)])
                self.geo64 = torch.as_tensor(geo, dtype=torch.float64, device=dev).contiguous()
        self.centers = torch.as_tensor(L.centers, dtype=torch.float64, device=dev).contiguous()
        self.complex_t = bool(L.T.is_complex()) if torch.is_tensor(L.T) else bool(np.iscomplexobj(L.T))
        if self.complex_t:
            Tt = torch.as_tensor(L.T).to(self.complex_dtype)
            self.T = torch.view_as_real(Tt).reshape(-1).to(dev).contiguous()
        else:
            self.T = torch.as_tensor(L.T, dtype=rd).to(dev).contiguous()
        self.leaf_slot = i32(L.leaf_slot)
        self.leaf_user = i32(L.leaf_user)

        self._stages = []

        def stage(S):
            t = {
                "box_off": i32(S.box_off), "box_bs": i32(S.box_bs), "box_rank": i32(S.box_rank),
                "box_toff": i64(S.box_toff), "gather_dst": i32(S.gather_dst), "gather_src": i32(S.gather_src),
                "items": i32(S.items.reshape(-1)),
            }
            self._stages.append(t)
            a = hip.LevelArgs()
            a.nboxes = int(S.box_off.shape[0])
            a.box_off = hip.ptr(t["box_off"])
            a.box_bs = hip.ptr(t["box_bs"])
            a.box_rank = hip.ptr(t["box_rank"])
            a.box_toff = hip.ptr(t["box_toff"])
            a.ngather = int(S.gather_dst.shape[0])
            a.gather_dst = hip.ptr(t["gather_dst"])
            a.gather_src = hip.ptr(t["gather_src"])
            a.nitems = int(S.items.shape[0])
            a.items = hip.ptr(t["items"])
            return a

        ups = [stage(S) for S in L.up]
        downs = [stage(S) for S in L.down]
        self._up_arr = (hip.LevelArgs * max(1, len(ups)))(*ups)
        self._down_arr = (hip.LevelArgs * max(1, len(downs)))(*downs)
        self._nup = len(ups)
        self._ndown = len(downs)

        self._nf = {k: i32(getattr(L, k)) for k in
                    ("g_pt", "g_n", "g_out", "g_box", "e_pt", "e_n", "e_q", "e_box", "e_flags")}
        self._nf["e_shift"] = self._entry_shifts(L)
        self.item_order = "heavy"
        self.rows_exclusive = self._rows_exclusive(L)
        self._nf["items"] = self._items_tensor(L)
        # symmetric split of the near field: pair lists + the one-sided rest (the tables as
        # uploaded; set_self_block re-derives the split from them)
        self._split_base = getattr(L, "split", None)
        self.self_block = 0
        self._upload_split(L)
        self.use_symmetric = self._split is not None
        self.sym_unroll = 0
        self.interactions = L.interactions
        self.nslots = L.nslots
        self.N = L.N
        # bound of every pair distance: the diameter of the points' bounding box
        P = np.asarray(L.points)
        self.max_distance = float(np.linalg.norm(P.max(axis=1) - P.min(axis=1))) if P.size else 0.0
        self._invalidate()
        return self

    def _invalidate(self):
        """Drop plans and cached graphs after a change of layout or kernel settings. Graphs
        already handed out (``capture``) keep their own buffers alive; ``generation`` tells a
        caller holding one that it no longer matches this object's settings."""
        for k in list(self._plans):
            self._destroy_plan(k)
        self._graphs = {}
        self.generation = getattr(self, "generation", 0) + 1

    def _entry_shifts(self, lists):
        """Precision form: per entry the shift of its points into its group's box frame (FP64
        difference of the box centres rounded once to FP32, as the kernels would form it), or None."""
        if self.precision != 1:
            return None
        C = np.asarray(self.layout.centers, dtype=np.float64)
        gbox = np.repeat(np.asarray(lists.g_box, np.int64), np.asarray(lists.g_e1) - np.asarray(lists.g_e0))
        sh = (C[gbox] - C[np.asarray(lists.e_box, np.int64)]).astype(np.float32)
        return torch.as_tensor(np.ascontiguousarray(sh.reshape(-1)), device=self.device)

    def _upload_split(self, L):
        """Device copies of the split tables of ``L`` (or none) and the interaction counts."""
        self._split = None
        if getattr(L, "split", None) is not None:
            self._split = []
            for part in L.split:
                t = {k: torch.as_tensor(np.ascontiguousarray(getattr(part, k), dtype=np.int32), device=self.device)
                     for k in ("g_pt", "g_n", "g_out", "g_box", "e_pt", "e_n", "e_q", "e_box", "e_flags")}
                t["e_shift"] = self._entry_shifts(part)
                t["items"] = self._items_tensor(part)
                self._split.append(t)
            self._split_structs = [hip.P2PLists(), hip.P2PLists()]
            self._fill_split_structs()
        self.evaluations = getattr(L, "evaluations", L.interactions)
        # pairs evaluated by the symmetric kernel / interactions of the one-sided rest
        self.sym_pairs = L.split[1].interactions if self._split is not None else 0
        self.one_interactions = L.split[0].interactions if self._split is not None else L.interactions

    def set_self_block(self, block):
        """Self segments (a target set against its own points) evaluated symmetrically in blocks of
        ``block`` targets: the pairs of different blocks join the symmetric kernel's tables, the
        diagonal blocks stay one-sided (``layout.split_self_blocks``). 0 (the default) keeps every
        self segment one-sided. Needs the symmetric split; the results agree to rounding."""
        from .layout import make_items, split_self_blocks

        block = int(block)
        if block < 0:
            raise ValueError(f"self block must be >= 0, got {block}")
        if self._split_base is None or block == self.self_block:
            return
        L = self.layout
        one, sym = split_self_blocks(*self._split_base, L.max_sources, block)
        for part in (one, sym):  # work items for the current source-run length
            part.items, part.item_cost = make_items(part.g_n, part.g_e0, part.g_e1, part.e_n, L.max_sources, part.tpi)
        L.split = (one, sym)
        L.evaluations = one.interactions + sym.interactions
        self.self_block = block
        self._upload_split(L)
        self._invalidate()

    def _fill_split_structs(self):
        for st, t in zip(self._split_structs, self._split):
            st.ngroups = int(t["g_pt"].shape[0])
            st.nentries = int(t["e_pt"].shape[0])
            st.nitems = int(t["items"].shape[0] // 4)
            for k in ("g_pt", "g_n", "g_out", "g_box", "e_pt", "e_n", "e_q", "e_box", "e_flags", "items", "e_shift"):
                setattr(st, k, hip.ptr(t[k]))

    def set_sym_unroll(self, n):
        """Sources per inner step of the symmetric near-field kernel (1, 2 or 4; 0 = default 2)."""
        self.sym_unroll = int(n)
        self._invalidate()

    def set_symmetric(self, on):
        """Use (True, default when available) or bypass the symmetric near-field kernel."""
        self.use_symmetric = bool(on) and self._split is not None
        self._invalidate()

    def _p2p_args(self, width, urev=None):
        a = hip.P2PArgs()
        a.kernel = self.operator.kernel_id
        a.precision = self.precision
        a.width = width
        a.kappa = self.kappa
        X = self.XX_slots
        a.px = hip.ptr(X[0])
        a.py = hip.ptr(X[1])
        a.pz = hip.ptr(X[2]) if self.ndim == 3 else None
        if self.normals is not None:
            a.nx = hip.ptr(self.normals[0])
            a.ny = hip.ptr(self.normals[1])
            a.nz = hip.ptr(self.normals[2])
        a.centers = hip.ptr(self.centers)
        if self.geo64 is not None:
            a.geo64 = hip.ptr(self.geo64)
            a.geo64_n = int(self.geo64.shape[1])
        nf = self._nf
        a.ngroups = int(nf["g_pt"].shape[0])
        for k in ("g_pt", "g_n", "g_out", "g_box", "e_pt", "e_n", "e_q", "e_box", "e_flags", "items", "e_shift"):
            setattr(a, k, hip.ptr(nf[k]))
        a.nentries = int(nf["e_pt"].shape[0])
        a.nitems = int(nf["items"].shape[0] // 4)
        a.mfma_min_channels = int(self.mfma_min_channels)
        # the apply zeroes the potentials before the near field, so single-writer rows are stored
        a.store_rows = 1 if self.rows_exclusive else 0
        a.sym_unroll = int(self.sym_unroll)
        a.max_distance = float(self.max_distance)
        uk = getattr(self.operator, "user_kernel", None)
        if uk is not None:  # user-defined operator: near field compiled at run time
            from .jit import compile_kernels

            for i, fn in enumerate(compile_kernels(uk, self.precision)):
                a.user_fn[i] = fn
        if self.use_symmetric and (self.precision == 0 or urev is not None):
            a.one = ctypes.pointer(self._split_structs[0])
            a.sym = ctypes.pointer(self._split_structs[1])
            if urev is not None:
                a.urev = hip.ptr(urev)
                a.urev_len = int(urev.numel())
        return a

    def _urev(self, width):
        """FP64 buffer for the symmetric kernel's reverse sums (FP32 form, one channel)."""
        cx = self.complex_vectors
        if not self.use_symmetric or self.precision != 1 or width != (2 if cx else 1):
            return None
        return torch.zeros(2 * self.nslots * width, dtype=torch.float64, device=self.device)

    def _plan(self, width):
        p = self._plans.get(width)
        if p is not None:
            return p
        rd = self.real_dtype
        qbuf = torch.zeros(2 * self.nslots * width, dtype=rd, device=self.device)
        ubuf = torch.zeros(2 * self.nslots * width, dtype=rd, device=self.device)
        d = hip.PlanDesc()
        d.precision = self.precision
        d.complex_t = 1 if self.complex_t else 0
        d.width = width
        d.nslots = self.nslots
        d.npoints = self.N
        d.leaf_slot = hip.ptr(self.leaf_slot)
        d.leaf_user = hip.ptr(self.leaf_user)
        d.T = hip.ptr(self.T)
        d.qbuf = hip.ptr(qbuf)
        d.ubuf = hip.ptr(ubuf)
        up_arr, down_arr = self._level_arrays(width)
        d.nup = self._nup
        d.up = ctypes.cast(up_arr, ctypes.POINTER(hip.LevelArgs))
        d.ndown = self._ndown
        d.down = ctypes.cast(down_arr, ctypes.POINTER(hip.LevelArgs))
        urev = self._urev(width)
        d.p2p = self._p2p_args(width, urev)
        handle = ctypes.c_void_p()
        hip.check(hip.lib().skf_plan_create(ctypes.byref(d), ctypes.byref(handle)))
        p = {"handle": handle, "qbuf": qbuf, "ubuf": ubuf, "urev": urev, "desc": d, "up_arr": up_arr,
             "down_arr": down_arr}
        self._plans[width] = p
        return p

    def _level_arrays(self, width):
        """Per-width copies of the level stages with their launch variants (``set_level_variants``)."""
        var = getattr(self, "_level_variants", {}).get(int(width), {})
        arrays = []
        for name, src, n in (("up", self._up_arr, self._nup), ("down", self._down_arr, self._ndown)):
            arr = (hip.LevelArgs * max(1, n))()
            vs = var.get(name, [])
            for i in range(n):
                arr[i] = hip.LevelArgs.from_buffer_copy(src[i])
                arr[i].variant = int(vs[i]) if i < len(vs) else 0
            arrays.append(arr)
        return arrays

    def set_level_variants(self, width, up=None, down=None):
        """Launch shape of every upward / downward stage for ``width`` channels (see
        ``skf_level_args.variant``; lists indexed like ``layout.up`` / ``layout.down``)."""
        if not hasattr(self, "_level_variants"):
            self._level_variants = {}
        v = self._level_variants.setdefault(int(width), {})
        if up is not None:
            v["up"] = [int(x) for x in up]
        if down is not None:
            v["down"] = [int(x) for x in down]
        self._invalidate()

    def level_variants(self, width):
        v = getattr(self, "_level_variants", {}).get(int(width), {})
        return {"up": list(v.get("up", [0] * self._nup)), "down": list(v.get("down", [0] * self._ndown))}

    def time_levels(self, width, reps=5):
        """Mean milliseconds of every upward and downward stage for ``width`` channels (HIP events on
        the current stream, the plan's own buffers): ({"up": [...], "down": [...]})."""
        p = self._plan(width)
        L = hip.lib()
        st = hip.stream_handle(self.device)
        es = p["qbuf"].element_size()
        half = self.nslots * width * es
        q0, u0 = p["qbuf"].data_ptr(), p["ubuf"].data_ptr()
        cx = 1 if self.complex_t else 0
        out = {}
        for name, arr, n, fn, a, b in (("up", p["up_arr"], self._nup, L.skf_up_level, q0, q0 + half),
                                       ("down", p["down_arr"], self._ndown, L.skf_down_level, u0, u0 + half)):
            times = []
            for i in range(n):
                call = lambda: hip.check(fn(ctypes.byref(arr[i]), hip.ptr(self.T), cx, ctypes.c_void_p(a),
                                            ctypes.c_void_p(b), width, self.precision, st))
                call()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    call()
                e1.record()
                torch.cuda.synchronize(self.device)
                times.append(e0.elapsed_time(e1) / reps)
            out[name] = times
        return out

    def _destroy_plan(self, width):
        p = self._plans.pop(width, None)
        if p is not None and p["handle"]:
            hip.lib().skf_plan_destroy(p["handle"])

    def __del__(self):
        try:
            for k in list(self._plans):
                self._destroy_plan(k)
        except Exception:
            pass

    # ------------------------------------------------------------------ apply
    @property
    def complex_vectors(self):
        """Charges and potentials are complex: complex kernels, and (as the reference casts them,
        ``batched_fmm.py:539-564``) every built-in kernel with kappa > 0; a user-defined real
        kernel keeps real vectors whatever its parameter."""
        return bool(self.operator.is_complex or (self.kappa > 0 and getattr(self.operator, "user_kernel", None) is None))

    def _vector_dtype(self, q):
        cplx = q.is_complex() or self.complex_vectors
        return self.complex_dtype if cplx else self.real_dtype

    def prepare_input(self, q):
        """Cast charges to the stored precision (real -> complex when kappa > 0) on the device."""
        dt = self._vector_dtype(q)
        return q.to(device=self.device, dtype=dt).contiguous()

    def apply_device(self, q, out=None):
        """``y = K~ q`` for a device tensor ``q`` of shape ``(N,)`` or ``(N, R)`` already in the apply dtype."""
        if q.shape[0] != self.N:
            raise ValueError("charge vector length must match the number of points")
        xr = torch.view_as_real(q) if q.is_complex() else q
        width = int(np.prod(xr.shape[1:])) if xr.dim() > 1 else 1
        p = self._plan(width)
        if out is None:
            out = torch.empty_like(q)
        yr = torch.view_as_real(out) if out.is_complex() else out
        # the registered PyTorch operator (C++ extension over skf_apply), on the current stream
        hip.torch_ops().apply(p["handle"].value, xr, yr)
        return out

    def matvec(self, q, verbose=False):
        """Fast apply; returns a tensor on ``q``'s device (reference: ``batched_fmm.py:539-564``)."""
        src_dev = q.device
        y = self.apply_device(self.prepare_input(q))
        return y if src_dev == self.device else y.to(src_dev)

    def forward(self, q):
        return self.matvec(q)

    def set_near_field_split(self, max_sources):
        """Re-split the near-field work items into source runs of at most ``max_sources`` points."""
        from .layout import resplit

        resplit(self.layout, max_sources)
        self._upload_items()

    def set_item_order(self, mode):
        """Launch order of the near-field work items: ``"heavy"`` (heaviest first, the default) or
        ``"xcd"`` (spatial runs per XCD, so that neighbouring items share the XCD's L2)."""
        if mode not in ITEM_ORDERS:
            raise ValueError(f"item order must be one of {ITEM_ORDERS}, got {mode!r}")
        self.item_order = mode
        self._upload_items()

    def _items_tensor(self, lists):
        items, _ = order_items(lists.items, lists.item_cost, self.item_order)
        return torch.as_tensor(np.ascontiguousarray(items.reshape(-1), dtype=np.int32), device=self.device)

    @staticmethod
    def _rows_exclusive(lists):
        """Every (group, 64-target block) has one work item, i.e. no group's sources are split
        over several items: each output row then has a single writer."""
        it = np.asarray(lists.items, dtype=np.int64).reshape(-1, 4)
        key = it[:, 0] * (1 << 31) + it[:, 1]
        return bool(np.unique(key).size == key.size)

    def _upload_items(self):
        self.rows_exclusive = self._rows_exclusive(self.layout)
        self._nf["items"] = self._items_tensor(self.layout)
        if self._split is not None:
            for t, part in zip(self._split, self.layout.split):
                t["items"] = self._items_tensor(part)
            self._fill_split_structs()
        self._invalidate()

    def time_near_field(self, width, reps=5, part="all"):
        """Mean duration (ms) of one near-field launch for ``width`` channels, HIP events on the current stream.

        ``part="sym"`` times the call with the one-sided rest emptied (zero fill, symmetric
        kernel, fold of the reverse sums), ``part="one"`` the rest alone; both only when the
        symmetric split is in use."""
        p = self._plan(width)
        a = self._p2p_args(width, p["urev"])
        if part != "all":
            if not (self.use_symmetric and a.sym):
                raise ValueError("near-field parts exist only with the symmetric split in use")
            empty = hip.P2PLists()
            keep = self._split_structs[1] if part == "sym" else self._split_structs[0]
            if part == "sym":
                a.one, a.sym = ctypes.pointer(empty), ctypes.pointer(keep)
            else:
                a.one, a.sym = ctypes.pointer(keep), ctypes.pointer(empty)
        a.q = hip.ptr(p["qbuf"])
        a.u = hip.ptr(p["ubuf"])
        st = hip.stream_handle(self.device)
        L = hip.lib()
        hip.check(L.skf_p2p(ctypes.byref(a), st))
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            hip.check(L.skf_p2p(ctypes.byref(a), st))
        e1.record()
        torch.cuda.synchronize(self.device)
        return e0.elapsed_time(e1) / reps

    def set_mfma_min_channels(self, n):
        """Channel count from which FP32 real kernels use the matrix-core near field (0 default 16, <0 never)."""
        self.mfma_min_channels = int(n)
        self._invalidate()

    # ------------------------------------------------------------ HIP graphs
    def _device_buffers(self):
        """Every device tensor an apply's kernels address (layout tables, points, T)."""
        keep = [self.T, self.XX_slots, self.normals, self.geo64, self.centers, self.leaf_slot, self.leaf_user]
        keep.append(dict(self._nf))
        keep.extend(dict(t) for t in (self._split or []))
        keep.extend(dict(t) for t in self._stages)
        return keep

    def capture(self, x, out):
        """Capture one apply ``x -> out`` (fixed device buffers) into a HIP graph.

        The apply issues only kernels (its zero fills are kernels, not memset nodes), so the
        graph replays the whole apply with one launch call: ``captured.replay()``. The returned
        ``CapturedApply`` holds the plan's work vectors and the layout tables the graph's
        kernels address, so a later setter (``set_symmetric``, ``set_near_field_split``, ...)
        or a new layout cannot free memory under it; ``captured.stale`` then turns True and
        a caller that wants the new settings captures again.
        """
        if x.shape[0] != self.N:
            raise ValueError("charge vector length must match the number of points")
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.apply_device(x, out=out)  # creates the plan and its work vectors outside the capture
        torch.cuda.current_stream(self.device).wait_stream(side)
        xr = torch.view_as_real(x) if x.is_complex() else x
        plan = self._plans[int(np.prod(xr.shape[1:])) if xr.dim() > 1 else 1]
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.apply_device(x, out=out)
        return CapturedApply(g, self, [plan, x, out] + self._device_buffers())

    # ------------------------------------------------------------ accounting
    def nbytes_proj(self) -> int:
        return int(self.T.numel() * self.T.element_size())

    def nbytes_lists(self) -> int:
        acc = self.XX_slots.numel() * self.XX_slots.element_size()
        if self.geo64 is not None:
            acc += self.geo64.numel() * 8
        for t in self._nf.values():
            acc += 0 if t is None else t.numel() * t.element_size()
        for part in self._split or []:
            for t in part.values():
                acc += 0 if t is None else t.numel() * t.element_size()
        for st in self._stages:
            for t in st.values():
                acc += t.numel() * t.element_size()
        acc += self.leaf_slot.numel() * 4 + self.leaf_user.numel() * 4
        return int(acc)


class CapturedApply:
    """A HIP graph of one apply together with every device buffer its kernels use.

    ``replay()`` launches the graph. ``stale`` is True once the apply object it came from
    has changed its layout or kernel settings (its ``generation`` moved on); the graph then
    still replays safely, with the settings it was captured with.
    """

    def __init__(self, graph, owner, buffers):
        self.graph = graph
        self.generation = owner.generation
        self._owner = owner
        self._buffers = buffers

    @property
    def stale(self):
        return self._owner.generation != self.generation

    def replay(self):
        self.graph.replay()


class BatchedInstances(SkelFMMBatched):
    """Independent point sets (same operator) applied together as one apply.

    The builds' layouts are merged (``layout.merge_layouts``): one near-field launch
    covers the interaction lists of every instance, and one upward and one downward
    stage per tree level covers every instance that has that level. Charges and
    results are the instances' arrays stacked along the point axis (``concat`` /
    ``split``); right-hand sides ride along the second axis.
    """

    def __init__(self, builds, use_single_precision=False, device=None, max_sources=16384):
        torch.nn.Module.__init__(self)
        if not builds:
            raise ValueError("need at least one build")
        b0 = builds[0]
        self.operator = operators_mod._require_operator(b0.operator)
        if self.operator.kernel_id < 0:
            raise NotImplementedError("the HIP apply supports the built-in operators of skelfmm_amd.operators only")
        for b in builds:
            if b.operator is not b0.operator or b.ndim != b0.ndim or float(b.kappa) != float(b0.kappa):
                raise ValueError("all instances must share operator, dimension and kappa")
        self.kappa = float(b0.kappa)
        self.ndim = b0.ndim
        self.kernel_backend = "hip"
        self._plans = {}
        self._graphs = {}
        self.generation = 0
        self.instance_sizes = [int(b.N) for b in builds]
        self.ninstances = len(builds)
        nf_dev = None
        if all(getattr(b, "build_device", "cpu") == "gpu" for b in builds):  # GPU builds: tables there too
            hip.require_gpu()
            nf_dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        L = merge_layouts([build_layout(b, max_sources=max_sources, near_field_device=nf_dev) for b in builds])
        self.upload_layout(L, use_single_precision, device)

    def concat(self, xs):
        """Stack per-instance charge arrays (N_i,) or (N_i, R) along the point axis on the device."""
        return torch.cat([torch.as_tensor(x) for x in xs], dim=0).to(self.device)

    def split(self, y):
        """Per-instance views of a stacked result."""
        return list(torch.split(y, self.instance_sizes, dim=0))
