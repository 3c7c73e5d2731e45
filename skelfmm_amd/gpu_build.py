"""GPU interpolative decompositions for the skeletonization build.

Two paths for the boxes of one tree level, both in FP64 (complex128 for Helmholtz):

* ``adaptive``: proxy matrices ``A_b`` (``skf_proxy_matrices``), column-pivoted
  Householder QR of each ``A_b`` stopped at the first ``|R_ii| <= tol |R_00|``
  (``skf_pqr_batched``) and ``T_b = R11^{-1} R12`` (``skf_trsm_batched``): the
  reference's adaptive ID (``skelfmm/util.py:13-51``) with early stopping instead
  of a full factorization.
* ``sketched`` (the learned-rank build): given a predicted rank ``k_b`` per box, a
  Gaussian sketch ``Y_b = Omega A_b`` with ``l_b = ceil(1.25 k_b) + 8`` rows (one
  batched GEMM) replaces ``A_b`` in the pivoted QR, which then runs on ``l_b`` rows
  instead of ``m``, with the cutoff tightened to ``tol / 2`` (a Gaussian sketch blurs
  it). A box is accepted when the QR of its sketch reaches the cutoff within ``l_b``
  steps and an independent 16-row sketch of the residual ``A[:, red] - A[:, skel] T``
  stays within ``6 tol`` of the largest sketched column norm (the proxy-residual
  check, ``skf_id_residual``); every other box falls back to the adaptive path.
  Boxes predicted nearly incompressible take the adaptive path directly, boxes
  predicted full rank keep all their points without a factorization.

Results are returned in the CPU build's convention: ``(rank, idx, proj)`` with
``idx`` the column pivots (host) and ``proj = T``, which stays in device memory
(``DeviceProj``) until the box layout gathers it there.
"""

from __future__ import annotations

import math
from time import perf_counter

import numpy as np
import torch

from . import hip
from .proxy import get_cube_surface, get_sphere_surface

__all__ = ["GPUSkeletonizer", "DeviceProj", "accept_sketch"]

LDS_BYTES = 64 * 1024


def _unit_proxy(operator, ndim, nproxy):
    """Proxy points of a box of side 1 / diameter 1 centred at the origin, as (P, 3)."""
    if operator.proxy_surface == "cube":
        u = get_cube_surface(np.zeros(ndim), 1.0, nproxy)
    else:
        u = get_sphere_surface(np.zeros(3), 1.0, nproxy)
    out = np.zeros((u.shape[0], 3))
    out[:, :u.shape[1]] = u
    return out


def accept_sketch(Z, result, lrows, tol, cap=None, accept=6.0):
    """Acceptance test of a sketched ID ``result = (rank, idx, T)`` from ``lrows`` sketch rows.

    Rejected when the sketch's QR ran out of rows before the rank cutoff
    (``rank >= lrows``: the predicted rank was too low -- unless ``lrows``
    already reaches ``cap = min(m, n)``, the largest possible rank), or when the residual
    ``Z[:, red] - Z[:, skel] T`` of an independent sketch ``Z`` of the proxy
    matrix has a column norm above ``accept * tol`` times the largest column norm
    of ``Z`` (the proxy-residual check at ``tol``). Host statement of the rule that
    ``GPUSkeletonizer.sketched`` applies to the residuals ``skf_id_residual`` returns.
    """
    rk, idx, T = result
    if rk >= lrows and (cap is None or lrows < cap):
        return False
    n = Z.shape[1]
    if n == 0 or rk == n:
        return True
    skel, red = Z[:, idx[:rk]], Z[:, idx[rk:]]
    E = red - skel @ T if rk > 0 else red
    den = float(np.max(np.linalg.norm(Z, axis=0)))
    return float(np.max(np.linalg.norm(E, axis=0))) <= accept * tol * max(den, 1e-300)


class DeviceProj:
    """Interpolation matrix ``T_b`` (``rank x (n - rank)``, row-major) left in device memory.

    ``buf`` is the flat FP64 result buffer of one batched TRSM launch (complex values as
    interleaved pairs) and ``off`` the box's first element in it. The box layout gathers
    these on the device (``layout.build_layout``); ``numpy.asarray(p)`` copies one matrix
    to the host for code that wants it there."""

    __slots__ = ("buf", "off", "shape", "dtype")

    def __init__(self, buf, off, shape, cplx):
        self.buf = buf
        self.off = int(off)
        self.shape = (int(shape[0]), int(shape[1]))
        self.dtype = np.dtype(np.complex128 if cplx else np.float64)

    @property
    def size(self):
        return self.shape[0] * self.shape[1]

    @property
    def nbytes(self):
        return self.size * self.dtype.itemsize

    @property
    def ndim(self):
        return 2

    def __array__(self, dtype=None, copy=None):
        w = 2 if self.dtype.kind == "c" else 1
        h = self.buf[self.off * w:(self.off + self.size) * w].cpu().numpy()
        a = (h.view(np.complex128) if w == 2 else h).reshape(self.shape)
        return a if dtype is None else a.astype(dtype)


class GPUSkeletonizer:
    """Batched GPU IDs for one build (operator, kappa, tolerance, points, normals)."""

    def __init__(self, operator, kappa, tol, points, geometry=None, device=None, budget_bytes=24e9, seed=0):
        hip.require_gpu()
        self.L = hip.lib()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.op = operator
        self.kappa = float(kappa)
        self.tol = float(tol)
        self.ndim = points.shape[1]
        P = np.zeros((points.shape[0], 3))
        P[:, :self.ndim] = points
        self.pts = torch.as_tensor(P, dtype=torch.float64, device=self.device).contiguous()
        self.nrm = None
        if geometry is not None:
            self.nrm = torch.as_tensor(np.asarray(geometry, dtype=np.float64), device=self.device).contiguous()
        self.stacked = not (operator.is_symmetric and geometry is None)
        self.cplx = bool(operator.is_complex)
        self.budget = float(budget_bytes)
        self.gen = torch.Generator(device=self.device)
        self.gen.manual_seed(seed)
        self.timings = {}
        self.sketch_stats = {}  # rejected sketches by reason: rows_exhausted (rank too low), residual
        self.profile = False  # True: synchronize and time every phase (proxy, gemm, qr, check, d2h, host)
        self.device_proj = True  # interpolation matrices stay on the device (DeviceProj)
        self._t = None

    # ----------------------------------------------------------------- common
    def _phase(self, name):
        """Close the running phase under ``name`` (only when profiling)."""
        if not self.profile:
            return
        torch.cuda.synchronize(self.device)
        now = perf_counter()
        if self._t is not None:
            self.timings[name] = self.timings.get(name, 0.0) + now - self._t
        self._t = now

    def _stream(self):
        return hip.stream_handle(self.device)

    def _chunks(self, ncols, rows):
        """Box index chunks whose matrices fit the byte budget."""
        es = 16 if self.cplx else 8
        out, cur, acc = [], [], 0
        for i, n in enumerate(ncols):
            b = rows * int(n) * es
            if cur and (acc + b > self.budget or len(cur) >= 32768):  # grid.y stays below 65536
                out.append(cur)
                cur, acc = [], 0
            cur.append(i)
            acc += b
        if cur:
            out.append(cur)
        return out

    def _proxy(self, cands, centers, length, nproxy, lda=0):
        """Proxy matrices of the given boxes; returns (A, a_off, m, unit count)."""
        unit = _unit_proxy(self.op, self.ndim, nproxy)
        P = unit.shape[0]
        m = 2 * P if self.stacked else P
        ncols = np.array([len(c) for c in cands], dtype=np.int32)
        idx = np.concatenate(cands).astype(np.int32)
        idx_off = np.concatenate(([0], np.cumsum(ncols)[:-1])).astype(np.int64)
        if lda > 0:
            a_off = (np.arange(len(cands), dtype=np.int64) * m * lda)
            total = len(cands) * m * lda
        else:
            a_off = np.concatenate(([0], np.cumsum(m * ncols.astype(np.int64))[:-1])).astype(np.int64)
            total = int(m * ncols.astype(np.int64).sum())
        dev = self.device
        A = torch.zeros(total * (2 if self.cplx else 1), dtype=torch.float64, device=dev)
        c3 = np.zeros((len(cands), 3))
        c3[:, :self.ndim] = centers
        t_idx = torch.as_tensor(idx, device=dev)
        t_ioff = torch.as_tensor(idx_off, device=dev)
        t_n = torch.as_tensor(ncols, device=dev)
        t_c = torch.as_tensor(c3, device=dev)
        t_side = torch.full((len(cands),), 2.95 * length, dtype=torch.float64, device=dev)
        t_unit = torch.as_tensor(unit, device=dev)
        t_aoff = torch.as_tensor(a_off, device=dev)
        hip.check_build(self.L.skf_proxy_matrices(
            self.op.kernel_id, self.ndim, self.kappa, hip.ptr(self.pts), hip.ptr(self.nrm), hip.ptr(t_idx),
            hip.ptr(t_ioff), hip.ptr(t_n), hip.ptr(t_c), hip.ptr(t_side), hip.ptr(t_unit), P,
            1 if self.stacked else 0, 1 if self.cplx else 0, hip.ptr(A), hip.ptr(t_aoff), int(lda), len(cands),
            int(ncols.max()), self._stream()))
        keep = (t_idx, t_ioff, t_n, t_c, t_side, t_unit)
        self._phase("proxy")
        return A, t_aoff, a_off, m, ncols, keep

    def _qr_trsm_dev(self, A, t_aoff, lda, rows, ncols, kmax, tol):
        """Batched pivoted QR (early stop at ``tol``) + triangular solves, results left on the device.

        ``lda``: per-box row strides (None: rows of ``ncols`` entries)."""
        dev = self.device
        nb = len(ncols)
        m_arr = np.asarray(rows, dtype=np.int32)
        ncols = np.asarray(ncols, dtype=np.int64)
        d = {"m": torch.as_tensor(m_arr, device=dev), "n": torch.as_tensor(ncols.astype(np.int32), device=dev),
             "k": torch.as_tensor(np.asarray(kmax, dtype=np.int32), device=dev)}
        d["lda"] = None if lda is None else torch.as_tensor(np.asarray(lda, dtype=np.int32), device=dev)
        p_off = np.concatenate(([0], np.cumsum(ncols)[:-1])).astype(np.int64)
        d["p_off"] = torch.as_tensor(p_off, device=dev)
        d["perm"] = torch.empty(max(1, int(ncols.sum())), dtype=torch.int32, device=dev)
        d["rdiag"] = torch.zeros(max(1, int(ncols.sum())), dtype=torch.float64, device=dev)
        d["rank"] = torch.empty(nb, dtype=torch.int32, device=dev)
        d["conv"] = torch.empty(nb, dtype=torch.int32, device=dev)
        max_m, max_n = int(m_arr.max()), int(ncols.max())
        hip.check_build(self.L.skf_pqr_batched(
            hip.ptr(A), hip.ptr(t_aoff), hip.ptr(d["lda"]), hip.ptr(d["m"]), hip.ptr(d["n"]), hip.ptr(d["k"]), float(tol),
            hip.ptr(d["rank"]), hip.ptr(d["conv"]), hip.ptr(d["perm"]), hip.ptr(d["p_off"]), hip.ptr(d["rdiag"]), nb,
            max_m, max_n, 1 if self.cplx else 0, self._stream()))
        self._phase("qr")
        r_h = d["rank"].cpu().numpy().astype(np.int64)
        c_h = d["conv"].cpu().numpy().astype(bool)
        nc = np.maximum(ncols - r_h, 0)
        sizes = np.where(c_h, r_h * nc, 0)
        t_off = np.concatenate(([0], np.cumsum(sizes)[:-1])).astype(np.int64)
        d["T"] = torch.zeros(max(1, int(sizes.sum())) * (2 if self.cplx else 1), dtype=torch.float64, device=dev)
        d["rank_solve"] = torch.as_tensor(np.where(c_h, r_h, 0).astype(np.int32), device=dev)
        d["t_off"] = torch.as_tensor(t_off, device=dev)
        hip.check_build(self.L.skf_trsm_batched(
            hip.ptr(A), hip.ptr(t_aoff), hip.ptr(d["lda"]), hip.ptr(d["n"]), hip.ptr(d["rank_solve"]), hip.ptr(d["T"]),
            hip.ptr(d["t_off"]), nb, max(1, int(nc.max()) if nc.size else 1), 1 if self.cplx else 0, self._stream()))
        host = {"rank": r_h, "conv": c_h, "p_off": p_off, "t_off": t_off, "ncols": ncols}
        self._phase("trsm")
        return d, host

    def _results(self, d, host, keep):
        """(rank, idx, proj) per box where ``keep`` holds, else None.

        ``idx`` (the column pivots) comes to the host; ``proj`` stays on the device as a
        ``DeviceProj`` (or a host array with ``device_proj=False``)."""
        perm_h = d["perm"].cpu().numpy().astype(np.int64)
        T_h = None
        if not self.device_proj:
            T_h = d["T"].cpu().numpy()
            if self.cplx:
                T_h = T_h.view(np.complex128)
        self._phase("d2h")
        out = []
        ncols, ranks, p_off, t_off = host["ncols"], host["rank"], host["p_off"], host["t_off"]
        empty = np.complex128 if self.cplx else np.float64
        for i in range(len(ncols)):
            if not keep[i]:
                out.append(None)
                continue
            n, r = int(ncols[i]), int(ranks[i])
            po, to = p_off[i], t_off[i]
            idx = perm_h[po:po + n]
            if r * (n - r) == 0:
                proj = np.zeros((r, n - r), dtype=empty)
            elif T_h is None:
                proj = DeviceProj(d["T"], to, (r, n - r), self.cplx)
            else:
                proj = T_h[to:to + r * (n - r)].reshape(r, n - r)
            out.append((r, idx, proj))
        self._phase("host")
        return out

    def _qr_trsm(self, A, t_aoff, lda, rows, ncols, kmax):
        d, host = self._qr_trsm_dev(A, t_aoff, None if lda == 0 else np.full(len(ncols), lda), rows, ncols, kmax,
                                    self.tol)
        return self._results(d, host, host["conv"])

    def _lds_ok(self, rows, n):
        es = 16 if self.cplx else 8
        return 16 * n + es * rows <= LDS_BYTES

    # ---------------------------------------------------------------- adaptive
    def adaptive(self, cands, centers, length, nproxy):
        """Adaptive IDs of a level's boxes (None where the QR workspace does not fit LDS)."""
        t0 = perf_counter()
        self._t = perf_counter()
        unitP = _unit_proxy(self.op, self.ndim, nproxy).shape[0]
        m = 2 * unitP if self.stacked else unitP
        results = [None] * len(cands)
        ok = [i for i, c in enumerate(cands) if self._lds_ok(m, len(c))]
        for ch in self._chunks([len(cands[i]) for i in ok], m):
            sel = [ok[i] for i in ch]
            A, t_aoff, _, m_, ncols, keep = self._proxy([cands[i] for i in sel], centers[sel], length, nproxy)
            kmax = np.minimum(m_, ncols)
            res = self._qr_trsm(A, t_aoff, 0, np.full(len(sel), m_), ncols, kmax)
            for i, r in zip(sel, res):
                results[i] = r
            del A, keep
        self.timings["adaptive"] = self.timings.get("adaptive", 0.0) + perf_counter() - t0
        return results

    # ---------------------------------------------------------------- sketched
    def sketched(self, cands, centers, length, nproxy, k_pred, oversample=1.25, extra=8, cut=0.5, accept=6.0,
                 check_rows=16, direct=0.9, keep_all=0.95):
        """Learned-rank IDs from Gaussian sketches; None where the box must fall back.

        Box b is sketched with ``l_b = min(m, n_b, ceil(oversample k_b) + extra)`` rows; the
        pivoted QR of the sketch stops at ``|R_ii| <= cut * tol * |R_00|`` (the sketch blurs the
        cutoff, a tighter one keeps the error at the adaptive ID's level). The ID is accepted
        when the QR stopped before running out of rows and the residual on an independent
        ``check_rows``-row sketch is within ``accept * tol`` (``accept_sketch``, computed on
        the GPU by ``skf_id_residual``). Boxes predicted to be nearly incompressible
        (``k_b >= direct * min(m, n_b)``: a sketch that size distorts too much to be
        accepted) are not sketched; they return ``"direct"`` and take the adaptive path. Boxes
        predicted full rank (``k_b >= keep_all * n_b`` with ``n_b <= m``) keep every candidate
        as skeleton (rank ``n_b``, empty ``T``: exact, no factorization) and return it tagged
        ``"kept"``; ``keep_all=None`` disables this.
        """
        t0 = perf_counter()
        self._t = perf_counter()
        unitP = _unit_proxy(self.op, self.ndim, nproxy).shape[0]
        m = 2 * unitP if self.stacked else unitP
        nb = len(cands)
        ncols_all = np.array([len(c) for c in cands], dtype=np.int64)
        want = np.ceil(oversample * np.asarray(k_pred, dtype=np.float64)).astype(np.int64) + extra
        lrows = np.maximum(np.minimum(np.minimum(ncols_all, m), want), 1)
        results = [None] * nb
        kp = np.asarray(k_pred, dtype=np.float64)
        ok = []
        for i in range(nb):
            n_i = int(ncols_all[i])
            if keep_all is not None and n_i <= m and kp[i] >= keep_all * n_i:
                results[i] = ("kept", (n_i, np.arange(n_i, dtype=np.int64),
                                       np.zeros((n_i, 0), dtype=np.complex128 if self.cplx else np.float64)))
            elif kp[i] >= direct * min(m, n_i):
                results[i] = "direct"
            elif self._lds_ok(m, int(ncols_all[i])):
                ok.append(i)
        # group boxes of similar width so the padded batch wastes little
        order = sorted(ok, key=lambda i: ncols_all[i])
        groups, cur = [], []
        es = 16 if self.cplx else 8
        for i in order:
            grow = (len(cur) + 1) * m * int(ncols_all[i]) * es
            if cur and (ncols_all[i] > 1.25 * ncols_all[cur[0]] + 16 or len(cur) >= 4096 or grow > self.budget):
                groups.append(cur)
                cur = []
            cur.append(i)
        if cur:
            groups.append(cur)
        cdt = torch.complex128 if self.cplx else torch.float64
        # batches of groups whose sketches fit the budget: one QR / TRSM / check launch per batch
        batches, cur, acc = [], [], 0
        for g in groups:
            b = len(g) * (int(lrows[g].max()) + check_rows) * int(ncols_all[g].max()) * es
            if cur and acc + b > self.budget:
                batches.append(cur)
                cur, acc = [], 0
            cur.append(g)
            acc += b
        if cur:
            batches.append(cur)
        for batch in batches:
            boxes = [i for g in batch for i in g]
            ysz = [len(g) * int(lrows[g].max()) * int(ncols_all[g].max()) for g in batch]
            zsz = [len(g) * check_rows * int(ncols_all[g].max()) for g in batch]
            Ybuf = torch.empty(sum(ysz), dtype=cdt, device=self.device)
            Zbuf = torch.empty(sum(zsz), dtype=cdt, device=self.device)
            y_off, z_off, lda = [], [], []
            yb = zb = 0
            for g, ys, zs in zip(batch, ysz, zsz):
                nmax = int(ncols_all[g].max())
                lmax = int(lrows[g].max())
                A, _, _, m_, _, keep = self._proxy([cands[i] for i in g], centers[g], length, nproxy, lda=nmax)
                Ab = (A.view(cdt) if self.cplx else A).view(len(g), m_, nmax)
                Om = torch.randn(lmax + check_rows, m_, dtype=torch.float64, device=self.device, generator=self.gen)
                Om = Om.to(cdt) if self.cplx else Om
                self._phase("host")
                torch.matmul(Om[:lmax], Ab, out=Ybuf[yb:yb + ys].view(len(g), lmax, nmax))
                torch.matmul(Om[lmax:], Ab, out=Zbuf[zb:zb + zs].view(len(g), check_rows, nmax))
                del A, Ab, keep
                self._phase("gemm")
                y_off.extend(yb + j * lmax * nmax for j in range(len(g)))
                z_off.extend(zb + j * check_rows * nmax for j in range(len(g)))
                lda.extend([nmax] * len(g))
                yb += ys
                zb += zs
            Yr = torch.view_as_real(Ybuf).reshape(-1) if self.cplx else Ybuf
            Zr = torch.view_as_real(Zbuf).reshape(-1) if self.cplx else Zbuf
            t_yoff = torch.as_tensor(np.asarray(y_off, dtype=np.int64), device=self.device)
            t_zoff = torch.as_tensor(np.asarray(z_off, dtype=np.int64), device=self.device)
            lg = lrows[boxes]
            ncb = ncols_all[boxes]
            d, host = self._qr_trsm_dev(Yr, t_yoff, lda, lg, ncb, lg, cut * self.tol)
            res = torch.empty(len(boxes), dtype=torch.float64, device=self.device)
            zn = torch.empty(len(boxes), dtype=torch.float64, device=self.device)
            hip.check_build(self.L.skf_id_residual(
                hip.ptr(Zr), hip.ptr(t_zoff), hip.ptr(d["lda"]), check_rows, hip.ptr(d["n"]), hip.ptr(d["rank_solve"]),
                hip.ptr(d["perm"]), hip.ptr(d["p_off"]), hip.ptr(d["T"]), hip.ptr(d["t_off"]), len(boxes),
                1 if self.cplx else 0, hip.ptr(res), hip.ptr(zn), self._stream()))
            res_h, zn_h = res.cpu().numpy(), zn.cpu().numpy()
            self._phase("check")
            cap = np.minimum(m, ncb)
            rows_ok = host["conv"] & ((host["rank"] < lg) | (lg >= cap))
            res_ok = res_h <= accept * self.tol * np.maximum(zn_h, 1e-300)
            good = rows_ok & res_ok
            st = self.sketch_stats
            st["rows_exhausted"] = st.get("rows_exhausted", 0) + int(np.sum(~rows_ok))
            st["residual"] = st.get("residual", 0) + int(np.sum(rows_ok & ~res_ok))
            for i, r in zip(boxes, self._results(d, host, good)):
                results[i] = r
            del Ybuf, Zbuf, d
        self.timings["sketched"] = self.timings.get("sketched", 0.0) + perf_counter() - t0
        return results
