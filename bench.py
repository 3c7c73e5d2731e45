"""Apply throughput of the skeletonized FMM on one MI355X.

Headline workload: 3D Laplace (LAPLACE_3D), N = 1,000,000 points i.i.d. uniform on [0,1]^3
(numpy default_rng(0)), charges RandomState(1).randn, tol 1e-6, leaf_size 400, one
right-hand side. A step is one apply (charges resident in HBM, result left in HBM), timed
as HIP graph replays. ``value`` is the FP32 precision form in Mpoints*RHS/s; the plain FP64
form, the FP32/FP64 speed-up, relative l2 errors of both forms against an FP64 direct sum at
1,000 sampled targets and the build seconds (host IDs, then GPU IDs) are reported beside it.

``roofline`` is the dominant kernel (the FP32 symmetric near field): algorithmic flops per
launch = interactions served x (f_op + 2 c R) (f_op flops per kernel value, c = 1 real / 4
complex, R right-hand sides; the symmetric kernel serves two interactions per evaluated
pair) over its mean launch time (HIP events on the stream it runs on), against the FP32
peak; the flops it executes (one kernel value per pair) are reported beside it.
``roofline.whole_apply`` bounds the whole apply: sum over phases of max(bytes / 8 TB/s,
flops / peak), over the measured apply time.
``cpu_baseline`` is the oracle's padded torch-CPU FP64 restatement of the reference apply,
run on every affinity core on two sampled fractions of every interaction list and
extrapolated to all pairs; ``cpu_baseline.validation`` checks that extrapolation against an
unsampled apply of the 100k-point 2D Laplace workload (tol 1e-5, leaf 200).

The same line carries the other workloads (``--only-main`` leaves them out), each applied
on its own: ``annuli_32x64rhs`` (32 independent curvy annuli, 2D Laplace, N_i log-uniform in
[1e4, 5e5], tol 1e-5, leaf 128, 64 float32 right-hand sides each, one graph-captured apply),
``mixture_2M`` (2M points of a 16-component Gaussian mixture, 3D Helmholtz kappa 20, tol
1e-6, leaf 256: adaptive GPU build vs learned-rank build, complex64 apply) and
``sphere_1M`` / ``sphere_1M_double_layer`` (1M points on the unit sphere, 3D Helmholtz
kappa 10 pi, tol 1e-5, leaf 256, normals = points; complex64 vs complex128).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--only-main] [--no-cpu] [--no-tune]
       python bench.py --workload annuli|sphere|mixture [--double-layer]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# flops per kernel value, accumulation excluded (FMA = 2): Laplace: dx, dy, dz (3), r^2 (5),
# rsqrt or log (1); Helmholtz 3D: + r = r^2 rsqrt (1), kappa r (1), sin, cos (2), e^{ikr}/r (2);
# double layer: + (x - y).n (5) and the (ikr - 1) r^-2 factor (4). Each interaction adds 2 c R
# for the K q accumulation (c = 1 real, 4 complex).
F_OP = {"laplace_3d": 9, "laplace_2d": 9, "helmholtz_3d": 15, "helmholtz_double_layer_3d": 24,
        "helmholtz_adjoint_double_layer_3d": 24}
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
HBM_BYTES_PER_S = 8.0e12
# HBM bytes per launch of the FP32 symmetric kernel on the headline workload, from two
# rocprofv3 --pmc passes (FETCH_SIZE x 2.00, WRITE_SIZE x 1.00: the factors a 512 MiB copy in the
# same run reads), keyed by (N, tol, leaf), with the setting that was profiled (symmetric unroll,
# source-run length, self block; the traffic barely depends on any of them);
# profiles/apply_laplace3d_uniform1M_hbm_traffic.md
P2P_FP32_TRAFFIC_BYTES = {(1_000_000, 1e-6, 400): {"bytes": 567_433_414 + 637_247_584, "sym_unroll": 2,
                                                   "max_sources": 4096, "self_block": 192}}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def affinity_cores():
    return len(os.sched_getaffinity(0))


def interaction_flops(name, cplx, nrhs):
    return F_OP[name] + 2 * (4 if cplx else 1) * nrhs


# ------------------------------------------------------------------------ roofline
def whole_apply_roofline(d, width, t_meas, name):
    """Sum over the apply's phases of max(bytes / BW, flops / peak), against the measured time.

    Bytes are compulsory traffic: per interpolation stage the T blocks (s_T r (bs - r) each) and
    the vectors (2 s width (bs + r)) plus its gathers; the charge gather and result scatter
    (2 s width N each, plus 8 N of indices); the zero fill of the potentials (2 s width nslots).
    The near field is counted by its flops (interactions x (f_op + 2 c R))."""
    L = d.layout
    s = 4 if d.precision else 8
    cx = d.complex_vectors
    nrhs = width // 2 if cx else width
    s_t = s * (2 if d.complex_t else 1)
    peak = PEAK_TFLOPS["fp32" if d.precision else "fp64"] * 1e12
    phases = {}

    def add(key, nbytes, flops=0.0):
        phases[key] = {"bytes": float(nbytes), "flops": float(flops),
                       "bound_ms": 1e3 * max(nbytes / HBM_BYTES_PER_S, flops / peak)}

    add("fill", 2 * s * width * d.nslots)
    add("gather_scatter", 2 * (2 * s * width + 8) * d.N)
    for direction in ("up", "down"):
        tb = vb = fl = gb = 0.0
        for S in getattr(L, direction):
            bs = np.asarray(S.box_bs, dtype=np.float64)
            rk = np.asarray(S.box_rank, dtype=np.float64)
            tb += s_t * float(np.sum(rk * (bs - rk)))
            vb += 2 * s * width * float(np.sum(bs + rk))
            fl += 2 * (4 if d.complex_t else 1) * nrhs * float(np.sum(rk * (bs - rk)))
            gb += (2 * s * width + 8) * float(np.asarray(S.gather_dst).shape[0])
        add(direction + "_interpolation", tb + vb + gb, fl)
    add("near_field", 0.0, float(d.interactions) * interaction_flops(name, cx, nrhs))
    bound = sum(p["bound_ms"] for p in phases.values())
    return {"bound_ms": round(bound, 4), "measured_ms": round(t_meas * 1e3, 4),
            "frac": round(bound / (t_meas * 1e3), 4),
            "phases": {k: {"bound_ms": round(v["bound_ms"], 4), "bytes": v["bytes"], "flops": v["flops"]}
                       for k, v in phases.items()},
            "rule": "sum over phases of max(bytes / 8 TB/s, flops / peak), over the measured apply"}


# ---------------------------------------------------------------------- CPU baseline
def cpu_apply_estimate(fmm, q, fraction):
    """Seconds of one full CPU apply of ``fmm``, from two sampled runs of the oracle's padded apply.

    Runs on ``fraction`` and ``2 fraction`` of the pairs of every list (interpolation passes in
    full both times), fits t = a + b f over the fraction f of pairs actually kept and evaluates
    it at f = 1. Returns (seconds, sample description)."""
    from oracle.batched_apply import BatchedApplyCPU

    pts = []
    for f in (fraction, 2 * fraction):
        ba = BatchedApplyCPU(fmm, bytes_budget=10e9, pair_fraction=f)
        if not pts:
            ba.matvec(q)  # warm-up
        tc = time.perf_counter()
        ba.matvec(q)
        pts.append((ba.pairs_kept / max(ba.pairs_total, 1), time.perf_counter() - tc))
    (f1, t1), (f2, t2) = pts
    b = max(t2 - t1, 0.0) / max(f2 - f1, 1e-12)
    a = max(t1 - b * f1, 0.0)
    desc = (f"padded torch-CPU restatement of the reference apply on {f1:.1%} and {f2:.1%} of the pairs of every "
            f"interaction list ({t1:.1f} / {t2:.1f} s) with full interpolation passes; linear fit t = a + b f "
            f"(a = {a:.2f} s) at f = 1")
    return a + b, desc


def cpu_baseline_validation(cores):
    """The extrapolation against an unsampled CPU apply of the 100k-point 2D Laplace workload
    (uniform [0,1]^2, default_rng(0); RandomState(1) charges; tol 1e-5, leaf 200)."""
    from oracle.batched_apply import BatchedApplyCPU
    from skelfmm_amd import operators
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.tree import BalancedTree

    N = 100_000
    X = np.random.default_rng(0).random((N, 2))
    q = np.random.RandomState(1).randn(N)
    tree = BalancedTree(X, leaf_size=200)
    f = RecursiveSkeletonization(tree, operators.LAPLACE_2D, kappa=0.0, tol=1e-5)
    f.skel_tree(npoints_max=500)
    f.setup_lists()
    est, desc = cpu_apply_estimate(f, q, 0.02)
    ba = BatchedApplyCPU(f, bytes_budget=10e9)
    ba.matvec(q)
    full = []
    for _ in range(2):
        tc = time.perf_counter()
        ba.matvec(q)
        full.append(time.perf_counter() - tc)
    out = {"workload": "2D Laplace, 100k uniform points, tol 1e-5, leaf 200", "full_apply_s": round(min(full), 3),
           "extrapolated_s": round(est, 3), "extrapolated_over_full": round(est / min(full), 3), "cores": cores,
           "full_apply_value": round(N / min(full) / 1e6, 6), "sample": desc}
    log(f"cpu baseline validation: {out}")
    return out


# ----------------------------------------------------------------- shared helpers
def _tuned_summary(t):
    """The autotuner's choice without its per-candidate timings."""
    if not t:
        return t
    out = {k: t[k] for k in ("max_sources", "sym_unroll", "self_block", "mfma_min_channels", "item_order",
                             "near_field_ms", "cached", "pinned")
           if k in t}
    lv = t.get("levels")
    if lv:
        out["levels"] = {k: lv[k] for k in ("up", "down", "levels_default_ms", "levels_tuned_ms")}
    return out


def graph_time(d, x, steps, warmup):
    """(seconds per graph replay of the whole apply, result)."""
    import torch

    out = torch.empty_like(x)
    g = d.capture(x, out)
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - ts) / steps, out


def eager_time(d, x, steps, warmup):
    import torch

    y = torch.empty_like(x)
    for _ in range(warmup):
        d.apply_device(x, out=y)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for _ in range(steps):
        d.apply_device(x, out=y)
    torch.cuda.synchronize()
    return (time.perf_counter() - ts) / steps, y


def rel_err(y, ref, idx):
    yh = y.cpu().numpy()[idx].astype(np.complex128 if np.iscomplexobj(ref) else np.float64)
    return float(np.linalg.norm(yh - ref) / np.linalg.norm(ref))


# ------------------------------------------------------------------------- annuli
def annuli_instances(n_inst=32, nrhs=64, seed=2):
    """32 curvy annuli with N_i log-uniform in [1e4, 5e5] and float32 N(0,1) charges (N_i, nrhs).

    rng = default_rng(seed): N_i = round(exp(U(ln 1e4, ln 5e5))) for all sets first, then per set
    theta ~ U[0, 2pi), u ~ U[0, 1), r = (1 + 0.25 cos 6 theta)(0.5 + 0.5 u), then the charges.
    """
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(1e4), np.log(5e5), n_inst))).astype(np.int64)
    pts, qs = [], []
    for n in ns:
        th = rng.uniform(0.0, 2 * np.pi, n)
        u = rng.uniform(0.0, 1.0, n)
        r = (1.0 + 0.25 * np.cos(6 * th)) * (0.5 + 0.5 * u)
        pts.append(np.stack([r * np.cos(th), r * np.sin(th)], axis=1))
        qs.append(rng.standard_normal((n, nrhs), dtype=np.float32))
    return ns, pts, qs


def run_annuli(args, steps, warmup, with_cpu):
    """Batched uneven 2D Laplace: every instance and RHS in one graph-captured apply."""
    import torch

    from oracle.direct import direct_sum_c
    from skelfmm_amd import autotune, hip, operators
    from skelfmm_amd.batched import BatchedInstances
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.tree import BalancedTree

    hip.require_gpu()
    op = operators.LAPLACE_2D
    tol, leaf = 1e-5, 128
    ns, pts, qs = annuli_instances(args.instances, args.nrhs)
    t0 = time.perf_counter()
    builds = []
    for P in pts:
        tree = BalancedTree(P, leaf_size=leaf)
        f = RecursiveSkeletonization(tree, op, kappa=0.0, tol=tol)
        f.skel_tree(npoints_max=500)
        f.setup_lists()
        builds.append(f)
    t_host = time.perf_counter() - t0
    d32 = BatchedInstances(builds, use_single_precision=True)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    d64 = BatchedInstances(builds, use_single_precision=False)
    ntot = int(ns.sum())
    width = args.nrhs
    log(f"annuli: {len(ns)} instances, {ntot} points, {width} RHS; build {build_s:.1f}s (host {t_host:.1f}); "
        f"interactions/apply {d32.interactions:.3e}")
    tuned = autotune.tune(d32, width) if not args.no_tune else None
    if tuned:
        log(f"autotune FP32: {tuned}")
    x32 = d32.concat([torch.as_tensor(q) for q in qs])
    x64 = x32.to(torch.float64)
    t32, y32 = graph_time(d32, x32, steps, warmup)
    t64, y64 = graph_time(d64, x64, steps, warmup)
    p32 = d32.time_near_field(width, reps=5) * 1e-3
    log(f"annuli apply (graph replay) FP32 {t32 * 1e3:.2f} ms, FP64 {t64 * 1e3:.2f} ms; near field FP32 "
        f"{p32 * 1e3:.2f} ms")

    # per-instance relative l2 error on 4 RHS columns at 100 sampled targets vs the FP64 direct sum
    errs32, errs64 = [], []
    for i, (P, q, a32, a64) in enumerate(zip(pts, qs, d32.split(y32), d64.split(y64))):
        idx = np.random.default_rng(i).choice(P.shape[0], min(100, P.shape[0]), replace=False)
        ref = direct_sum_c(op.name, P[idx], P, q[:, :4].astype(np.float64))
        errs32.append(float(np.linalg.norm(a32[idx, :4].cpu().numpy().astype(np.float64) - ref) / np.linalg.norm(ref)))
        errs64.append(float(np.linalg.norm(a64[idx, :4].cpu().numpy() - ref) / np.linalg.norm(ref)))
    log(f"annuli rel-l2 vs direct sum (max over instances): FP32 {max(errs32):.3e}, FP64 {max(errs64):.3e}")

    # the same instances built on the GPU (tree, batched adaptive IDs, device work tables)
    gpu_build = None
    if not args.no_gpu_build:
        from skelfmm_amd.gpu_tree import GPUBalancedTree

        # first-call costs (rocPRIM, the device-table kernels) outside the timing
        wt = GPUBalancedTree(pts[0][:20000], leaf_size=leaf)
        wf = RecursiveSkeletonization(wt, op, kappa=0.0, tol=tol)
        wf.skel_tree(npoints_max=500, device="gpu")
        wf.setup_lists()
        BatchedInstances([wf], use_single_precision=True)
        del wt, wf
        torch.cuda.synchronize()
        tg = time.perf_counter()
        gbuilds = []
        for P in pts:
            gt = GPUBalancedTree(P, leaf_size=leaf)
            gf = RecursiveSkeletonization(gt, op, kappa=0.0, tol=tol)
            gf.skel_tree(npoints_max=500, device="gpu")
            gf.setup_lists()
            gbuilds.append(gf)
        dg = BatchedInstances(gbuilds, use_single_precision=True)
        torch.cuda.synchronize()
        t_g = time.perf_counter() - tg
        yg = dg.matvec(x32)
        errs_g = []
        for i, (P, q, a) in enumerate(zip(pts, qs, dg.split(yg))):
            idx = np.random.default_rng(i).choice(P.shape[0], min(100, P.shape[0]), replace=False)
            ref = direct_sum_c(op.name, P[idx], P, q[:, :4].astype(np.float64))
            errs_g.append(float(np.linalg.norm(a[idx, :4].cpu().numpy().astype(np.float64) - ref) / np.linalg.norm(ref)))
        gpu_build = {"build_s": round(t_g, 2), "host_build_s": round(build_s, 2),
                     "rank_sum": int(sum(int(b.rank_list.sum()) for b in gbuilds)),
                     "host_rank_sum": int(sum(int(b.rank_list.sum()) for b in builds)),
                     "fp32_max_rel_l2_err": max(errs_g)}
        log(f"annuli GPU build: {t_g:.2f}s (host {build_s:.1f}s), ranks {gpu_build['rank_sum']} "
            f"(host {gpu_build['host_rank_sum']}), FP32 apply error {max(errs_g):.3e}")
        del dg, yg, gbuilds

    mfma_tf = 2.0 * d32.interactions * width / p32 / 1e12
    out = {
        "value": round(ntot * width / t32 / 1e6, 3), "unit": "Mpoints*RHS/s", "ms_per_step": round(t32 * 1e3, 4),
        "dtype": "f32", "steps": steps, "warmup": warmup,
        "data": "synthetic: 32 curvy annuli, N_i log-uniform in [1e4, 5e5] (default_rng(2)), float32 N(0,1) charges",
        "config": {"workload": "batched_annuli_laplace_2d", "instances": int(len(ns)), "N_total": ntot,
                   "nrhs": width, "tol": tol, "leaf_size": leaf, "graph": True,
                   "interactions_per_apply": int(d32.interactions), "autotune": _tuned_summary(tuned)},
        "fp64": {"value": round(ntot * width / t64 / 1e6, 3), "ms_per_step": round(t64 * 1e3, 4),
                 "max_rel_l2_err": max(errs64)},
        "fp32_max_rel_l2_err": max(errs32),
        "build_s": round(build_s, 2),
        "gpu_build": gpu_build,
        "roofline": {"bound": "mfma", "achieved": round(mfma_tf, 3), "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                     "frac": round(mfma_tf / PEAK_TFLOPS["fp32"], 4),
                     "kernel": "near field (matrix-core kernel at >= 16 RHS unless the autotuner picked VALU)",
                     "kernel_ms": round(p32 * 1e3, 4), "flops_counted": "2 x interactions x RHS (K q products)",
                     "whole_apply": whole_apply_roofline(d32, width, t32, op.name)},
    }
    if with_cpu:
        import torch as _t

        from oracle.batched_apply import BatchedApplyCPU

        ncores = affinity_cores()
        _t.set_num_threads(ncores)
        k = int(np.argmin(ns))
        ba = BatchedApplyCPU(builds[k], bytes_budget=10e9)
        q0 = qs[k].astype(np.float64)
        ba.matvec(q0)
        tc = time.perf_counter()
        ba.matvec(q0)
        el = time.perf_counter() - tc
        out["cpu_baseline"] = {
            "value": round(int(ns[k]) * width / el / 1e6, 6), "unit": "Mpoints*RHS/s", "cores": ncores, "kind": "port",
            "sample": f"padded torch-CPU FP64 restatement of the reference apply, {width} RHS, whole apply of the "
                      f"smallest instance (N={int(ns[k])}, {el:.1f}s)"}
    return out


# ---------------------------------------------------------------------- Helmholtz
def mixture_points(n=2_000_000, seed=3, ncomp=16, sigma=0.02):
    """Isotropic Gaussian mixture in [0,1]^3: default_rng(seed) draws the centres U[0,1]^3,
    then Dirichlet(1) weights, then the component of every point, then the offsets; points
    outside the cube are clipped to it (the count is returned)."""
    rng = np.random.default_rng(seed)
    c = rng.random((ncomp, 3))
    w = rng.dirichlet(np.ones(ncomp))
    comp = rng.choice(ncomp, size=n, p=w)
    x = c[comp] + sigma * rng.standard_normal((n, 3))
    clipped = int(np.sum(np.any((x < 0) | (x > 1), axis=1)))
    return np.clip(x, 0.0, 1.0), clipped


def run_helmholtz(args, workload, double_layer, steps, warmup, with_cpu, n=None):
    """3D Helmholtz: ``sphere`` (1M points on the unit sphere, kappa 10 pi, tol 1e-5, leaf 256,
    single or double layer, adaptive GPU build) or ``mixture`` (2M-point Gaussian mixture, kappa
    20, tol 1e-6, leaf 256, built twice on the GPU: adaptive IDs, then learned-rank sketched IDs
    with adaptive fallback). Value: complex64 apply throughput."""
    import torch

    from oracle.direct import direct_sum_c
    from oracle.kernels import GEOM_SIDE
    from skelfmm_amd import hip, learned_rank, operators
    from skelfmm_amd.batched import SkelFMMBatched
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.gpu_tree import GPUBalancedTree

    hip.require_gpu()
    dev = torch.device("cuda", 0)
    extra = {}
    if workload == "sphere":
        N = n or 1_000_000
        v = np.random.default_rng(0).normal(size=(N, 3))
        X = v / np.linalg.norm(v, axis=1, keepdims=True)
        op = operators.HELMHOLTZ_DOUBLE_LAYER_3D if double_layer else operators.HELMHOLTZ_3D
        geom = X.copy() if double_layer else None
        kappa, tol, leaf = 10 * np.pi, 1e-5, 256
        name = "helmholtz_3d_sphere" + ("_double_layer" if double_layer else "")
        data = "synthetic: unit-sphere points default_rng(0) normal/|.|, normals = points"
    else:
        N = n or 2_000_000
        X, clipped = mixture_points(N)
        op, geom = operators.HELMHOLTZ_3D, None
        kappa, tol, leaf = 20.0, 1e-6, 256
        name = "helmholtz_3d_gaussian_mixture"
        data = f"synthetic: 16-component Gaussian mixture default_rng(3), sigma 0.02, {clipped} points clipped to [0,1]^3"
        extra["clipped_points"] = clipped
    rs = np.random.RandomState(1)
    q = rs.randn(N) + 1j * rs.randn(N)

    GPUBalancedTree(X[:4096], leaf_size=leaf)  # first-call costs (library load, rocPRIM) outside the timing
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree = GPUBalancedTree(X, leaf_size=leaf)
    torch.cuda.synchronize()
    t_tree = time.perf_counter() - t0
    log(f"{name}: N={N}, GPU tree {t_tree:.2f}s, levels {tree.nlevels}, boxes {tree.nboxes}")

    def build(rank_model):
        f = RecursiveSkeletonization(tree, op, kappa=kappa, tol=tol, geometry=geom)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        f.skel_tree(npoints_max=500, device="gpu", rank_model=rank_model)
        torch.cuda.synchronize()
        t_id = time.perf_counter() - ts
        ts = time.perf_counter()
        f.setup_lists()
        d = SkelFMMBatched(tree, op, kappa=kappa).build_from_copy(f, use_single_precision=False)
        torch.cuda.synchronize()
        return f, d, t_id, time.perf_counter() - ts

    builds = {"tree_s": round(t_tree, 3), "tree": "GPU (HIP kernels; identical to the host tree)"}
    f_a, d_a, tid_a, tset_a = build(None)
    builds["adaptive"] = {"id_s": round(tid_a, 3), "lists_upload_s": round(tset_a, 2),
                          "build_s": round(t_tree + tid_a + tset_a, 2), "rank_sum": int(f_a.rank_list.sum()),
                          "level_s": {int(k): round(v, 3) for k, v in f_a.level_times.items()}}
    log(f"adaptive GPU build: IDs {tid_a:.2f}s, lists+upload {tset_a:.2f}s, root level {f_a.root_level}")
    f_use, d64 = f_a, d_a
    model = learned_rank.load_default() if workload == "mixture" else None
    if model is not None:
        f_l, d_l, tid_l, tset_l = build(model)
        st = f_l.gpu_stats
        builds["learned"] = {"id_s": round(tid_l, 3), "lists_upload_s": round(tset_l, 2),
                             "build_s": round(t_tree + tid_l + tset_l, 2), "rank_sum": int(f_l.rank_list.sum()),
                             "stats": st, "fallback_rate": round(st["fallback"] / max(1, st["sketched"] + st["fallback"]), 4),
                             "rejected_by": dict(f_l.gpu_skel.sketch_stats),
                             "level_s": {int(k): round(v, 3) for k, v in f_l.level_times.items()}}
        builds["id_speedup_learned_over_adaptive"] = round(tid_a / tid_l, 3)
        builds["build_speedup_learned_over_adaptive"] = round(
            builds["adaptive"]["build_s"] / builds["learned"]["build_s"], 3)
        log(f"learned GPU build: IDs {tid_l:.2f}s ({tid_a / tid_l:.2f}x), whole build "
            f"{builds['build_speedup_learned_over_adaptive']:.2f}x, stats {st}")
        f_use, d64 = f_l, d_l
    d32 = SkelFMMBatched(tree, op, kappa=kappa).build_from_copy(f_use, use_single_precision=True)
    tuned = None
    if not args.no_tune:
        from skelfmm_amd import autotune

        tuned = {"complex64": _tuned_summary(autotune.tune(d32, 2)), "complex128": _tuned_summary(autotune.tune(d64, 2))}
        log(f"{name} autotune: {tuned}")

    q128 = torch.as_tensor(q, dtype=torch.complex128, device=dev)
    q64 = q128.to(torch.complex64)
    t64, y64 = graph_time(d32, q64, steps, warmup)
    t128, y128 = graph_time(d64, q128, steps, warmup)
    p64 = d32.time_near_field(2, reps=5) * 1e-3
    p128 = d64.time_near_field(2, reps=5) * 1e-3
    log(f"{name} apply (graph replay) complex64 {t64 * 1e3:.2f} ms, complex128 {t128 * 1e3:.2f} ms; near field "
        f"{p64 * 1e3:.2f} / {p128 * 1e3:.2f} ms")

    idx = np.random.default_rng(0).choice(N, min(args.ncheck, N), replace=False)
    side = GEOM_SIDE.get(op.name)
    ref = direct_sum_c(op.name, X[idx], X, q, kappa=kappa, geom_trg=geom[idx] if side == "target" else None,
                       geom_src=geom if side == "source" else None)
    errs = {"complex64": rel_err(y64, ref, idx), "complex128": rel_err(y128, ref, idx)}
    if "learned" in builds:
        _, ya = eager_time(d_a, q128, 1, 0)
        errs["complex128_adaptive_build"] = rel_err(ya, ref, idx)
    log(f"{name} rel-l2 vs direct sum at {idx.size} targets: {errs}")

    fpi = interaction_flops(op.name, True, 1)
    tf = d32.interactions * fpi / p64 / 1e12
    out = {
        "value": round(N / t64 / 1e6, 3), "unit": "Mpoints*RHS/s", "ms_per_step": round(t64 * 1e3, 4),
        "dtype": "c64", "steps": steps, "warmup": warmup,
        "data": data + "; charges RandomState(1) randn + 1j randn",
        "config": {"workload": name, "N": N, "kappa": kappa, "tol": tol, "leaf_size": leaf, "nrhs": 1,
                   "nlevels": int(tree.nlevels), "root_level": int(f_use.root_level),
                   "interactions_per_apply": int(d32.interactions), "autotune": tuned, **extra},
        "complex128": {"value": round(N / t128 / 1e6, 3), "ms_per_step": round(t128 * 1e3, 4)},
        "rel_l2_err": errs,
        "builds": builds,
        "roofline": {"bound": "valu", "achieved": round(tf, 3), "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                     "frac": round(tf / PEAK_TFLOPS["fp32"], 4),
                     "kernel": f"near field complex64 ({op.name}), all lists",
                     "kernel_ms": round(p64 * 1e3, 4), "flops_per_interaction": fpi,
                     "whole_apply": whole_apply_roofline(d32, 2, t64, op.name)},
    }
    if with_cpu:
        import torch as _t

        ncores = affinity_cores()
        _t.set_num_threads(ncores)
        est, desc = cpu_apply_estimate(f_use, q, 0.01)
        out["cpu_baseline"] = {"value": round(N / est / 1e6, 6), "unit": "Mpoints*RHS/s", "cores": ncores,
                               "kind": "port", "sample": "complex128 " + desc}
    return out


# --------------------------------------------------------------- headline workload
def run_cube(args):
    import torch

    from skelfmm_amd import hip, operators
    from skelfmm_amd.batched import SkelFMMBatched
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.tree import BalancedTree

    hip.lib()  # compiles first when missing or older than its sources
    hip.require_gpu()
    dev = torch.device("cuda", 0)
    if args.copy_probe:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copy probe: 536870912 bytes read + 536870912 bytes written")

    N = args.n
    X = np.random.default_rng(0).random((N, 3))
    q = np.random.RandomState(1).randn(N)
    op = operators.LAPLACE_3D

    t0 = time.perf_counter()
    tree = BalancedTree(X, leaf_size=args.leaf)
    t_tree = time.perf_counter() - t0
    t1 = time.perf_counter()
    fmm = RecursiveSkeletonization(tree, op, kappa=0.0, tol=args.tol)
    fmm.skel_tree(npoints_max=500)
    t_skel = time.perf_counter() - t1
    t2 = time.perf_counter()
    fmm.setup_lists()
    d64 = SkelFMMBatched(tree, op).build_from_copy(fmm, use_single_precision=False)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t2
    build_s = time.perf_counter() - t0
    d32 = SkelFMMBatched(tree, op).build_from_copy(fmm, use_single_precision=True)
    log(f"build {build_s:.1f}s (tree {t_tree:.1f}, skel {t_skel:.1f}, lists+upload {t_setup:.1f}); "
        f"levels {tree.nlevels}, root {fmm.root_level}, max rank {fmm.max_rank}, "
        f"interactions/apply {d64.interactions:.3e}")

    tuned = None
    if args.max_sources or args.sym_unroll or args.self_block:  # a pinned setting (profiling runs)
        for d in (d64, d32):
            if args.max_sources:
                d.set_near_field_split(args.max_sources)
            if args.sym_unroll:
                d.set_sym_unroll(args.sym_unroll)
            if args.self_block:
                d.set_self_block(args.self_block)
        tuned = {k: {"max_sources": int(d.layout.max_sources), "sym_unroll": int(d.sym_unroll or 2),
                     "self_block": int(d.self_block), "pinned": True}
                 for k, d in (("fp64", d64), ("fp32", d32))}
    elif not args.no_tune:
        from skelfmm_amd import autotune

        tuned = {"fp64": autotune.tune(d64, 1), "fp32": autotune.tune(d32, 1)}
        log(f"autotune: {tuned}")
    q64 = torch.as_tensor(q, dtype=torch.float64, device=dev)
    q32 = q64.to(torch.float32)

    e64_t, _ = eager_time(d64, q64, args.steps, args.warmup)
    t64, y64 = graph_time(d64, q64, args.steps, args.warmup)
    e32_t, _ = eager_time(d32, q32, args.steps, args.warmup)
    t32, y32 = graph_time(d32, q32, args.steps, args.warmup)
    log(f"apply (graph replay) FP64 {t64 * 1e3:.2f} ms, FP32 {t32 * 1e3:.2f} ms, ratio {t64 / t32:.2f}; "
        f"eager FP64 {e64_t * 1e3:.2f} ms, FP32 {e32_t * 1e3:.2f} ms")

    # the drop-in call: host numpy charges in, host numpy potentials out (PCIe copies included)
    qh32 = q.astype(np.float32)
    d32.matvec(torch.as_tensor(qh32)).numpy()
    ts = time.perf_counter()
    for _ in range(3):
        d32.matvec(torch.as_tensor(qh32)).numpy()
    t_host32 = (time.perf_counter() - ts) / 3
    log(f"host-buffer apply FP32 (numpy in/out): {t_host32 * 1e3:.2f} ms")

    # near-field launch alone and its symmetric part (zero fill + p2p_sym_kernel + fold), HIP
    # events on the stream it runs on
    p64 = d64.time_near_field(1, reps=10) * 1e-3
    p32 = d32.time_near_field(1, reps=10) * 1e-3
    s32 = d32.time_near_field(1, reps=10, part="sym") * 1e-3 if d32.use_symmetric else None
    s64 = d64.time_near_field(1, reps=10, part="sym") * 1e-3 if d64.use_symmetric else None
    fpi = interaction_flops(op.name, False, 1)
    log(f"near field FP64 {p64 * 1e3:.2f} ms, FP32 {p32 * 1e3:.2f} ms; symmetric part FP64 "
        f"{(s64 or 0) * 1e3:.2f} ms, FP32 {(s32 or 0) * 1e3:.2f} ms")

    from oracle.direct import direct_sum_c

    idx = np.random.default_rng(0).choice(N, min(args.ncheck, N), replace=False)
    ref = direct_sum_c(op.name, X[idx], X, q)
    e64, e32 = rel_err(y64, ref, idx), rel_err(y32, ref, idx)
    log(f"rel-l2 vs direct sum at {idx.size} targets: FP64 {e64:.3e}, FP32 {e32:.3e}")

    # the same tree skeletonized with the IDs on the GPU (batched adaptive pivoted QR)
    gpu_build = None
    if not args.no_gpu_build:
        from skelfmm_amd.gpu_tree import GPUBalancedTree

        # first-call costs (rocPRIM, the device-table kernels) outside the timing: a 20k-point build
        wt = GPUBalancedTree(X[:20000], leaf_size=args.leaf)
        wf = RecursiveSkeletonization(wt, op, kappa=0.0, tol=args.tol)
        wf.skel_tree(npoints_max=500, device="gpu")
        wf.setup_lists()
        SkelFMMBatched(wt, op).build_from_copy(wf, use_single_precision=True)
        del wt, wf
        torch.cuda.synchronize()
        tg = time.perf_counter()
        gtree = GPUBalancedTree(X, leaf_size=args.leaf)
        torch.cuda.synchronize()
        t_gtree = time.perf_counter() - tg
        tg = time.perf_counter()
        fg = RecursiveSkeletonization(gtree, op, kappa=0.0, tol=args.tol)
        fg.skel_tree(npoints_max=500, device="gpu")
        torch.cuda.synchronize()
        t_gid = time.perf_counter() - tg
        tg = time.perf_counter()
        fg.setup_lists()
        dg = SkelFMMBatched(gtree, op).build_from_copy(fg, use_single_precision=True)
        torch.cuda.synchronize()
        t_gset = time.perf_counter() - tg
        eg = rel_err(dg.matvec(q32), ref, idx)
        gpu_build = {"tree_s": round(t_gtree, 3), "id_s": round(t_gid, 3), "lists_upload_s": round(t_gset, 3),
                     "build_s": round(t_gtree + t_gid + t_gset, 2), "rank_sum": int(fg.rank_list.sum()),
                     "host_rank_sum": int(fmm.rank_list.sum()), "fp32_rel_l2_err": eg}
        log(f"GPU build: tree {t_gtree:.2f}s, IDs {t_gid:.2f}s, lists+upload {t_gset:.2f}s, FP32 apply error {eg:.3e}")
        del dg, fg

    cpu = None
    if not args.no_cpu:
        ncores = affinity_cores()
        torch.set_num_threads(ncores)
        est, desc = cpu_apply_estimate(fmm, q, args.cpu_fraction)
        cpu = {"value": round(N / est / 1e6, 6), "unit": "Mpoints*RHS/s", "cores": ncores, "kind": "port",
               "sample": "FP64 " + desc, "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
        cpu["validation"] = cpu_baseline_validation(ncores)
        log(f"cpu baseline: {cpu}")

    sym = bool(d32.use_symmetric and s32)
    served = 2 * d32.sym_pairs if sym else d32.interactions
    k_ms = s32 if sym else p32
    ach = served * fpi / k_ms / 1e12
    executed = (d32.sym_pairs * (fpi + 2) if sym else d32.interactions * fpi) / k_ms / 1e12
    setting = (int(d32.sym_unroll or 2), int(d32.layout.max_sources), int(d32.self_block))
    profiled = P2P_FP32_TRAFFIC_BYTES.get((N, args.tol, args.leaf)) if sym else None
    roofline = {
        "bound": "valu",
        "achieved": round(ach, 3),
        "peak": PEAK_TFLOPS["fp32"],
        "unit": "TFLOP/s",
        "frac": round(ach / PEAK_TFLOPS["fp32"], 4),
        "traffic": profiled["bytes"] if profiled else None,
        "traffic_unit": "bytes per launch (HBM read + write) of the dominant kernel, rocprofv3 FETCH_SIZE / "
                        "WRITE_SIZE scaled by a 512 MiB copy",
        "traffic_profiled_setting": None if not profiled else {k: profiled[k] for k in ("sym_unroll", "max_sources",
                                                                                        "self_block")},
        "run_setting": {"sym_unroll": setting[0], "max_sources": setting[1], "self_block": setting[2]},
        "kernel": ("p2p_sym_kernel<float, LAPLACE_3D> (pairs of lists, both directions; zero fill and fold of the "
                   "reverse sums included)" if sym else "p2p_kernel<float, LAPLACE_3D>"),
        "kernel_ms": round(k_ms * 1e3, 4),
        "flops_counted": f"interactions served x (f_op {F_OP[op.name]} + 2 c R) = {served} x {fpi}",
        "executed_tflops": round(executed, 3),
        "executed_counted": f"one kernel value per evaluated pair: pairs x (f_op + 2 x 2) = {d32.sym_pairs} x {fpi + 2}",
        "peak_note": "157.3 TFLOP/s is the published FP32 rate (packed FP32); this build keeps packed FP32 off, "
                     "so the non-packed issue limit is 78.6 TFLOP/s",
        "fp64_kernel": {"achieved": round(served * fpi / (s64 if (sym and s64) else p64) / 1e12, 3),
                        "peak": PEAK_TFLOPS["fp64"], "kernel_ms": round((s64 if (sym and s64) else p64) * 1e3, 4),
                        "near_field_call_ms": round(p64 * 1e3, 4)},
        "near_field_call_ms": {"fp32": round(p32 * 1e3, 4), "fp64": round(p64 * 1e3, 4)},
        "whole_apply": whole_apply_roofline(d32, 1, t32, op.name),
        "whole_apply_fp64": whole_apply_roofline(d64, 1, t64, op.name),
    }
    from skelfmm_amd import hip as _hip

    return {
        "metric": "Mpoints*RHS/s per apply, FP32 precision form (FP64 form and build s beside it)",
        "value": round(N / t32 / 1e6, 3),
        "unit": "Mpoints*RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: uniform [0,1]^3 points default_rng(0), RandomState(1) N(0,1) charges",
        "config": {"workload": "laplace_3d_uniform_cube", "N": N, "tol": args.tol, "leaf_size": args.leaf,
                   "nrhs": 1, "root_level": int(fmm.root_level), "nlevels": int(tree.nlevels),
                   "interactions_per_apply": int(d64.interactions),
                   "autotune": None if tuned is None else {k: _tuned_summary(v) for k, v in tuned.items()}},
        "build_id": _hip.build_id(),
        "fp64": {"value": round(N / t64 / 1e6, 3), "ms_per_step": round(t64 * 1e3, 4), "rel_l2_err": e64,
                 "eager_ms_per_step": round(e64_t * 1e3, 4)},
        "timing": "HIP graph replays of the whole apply (charges and result resident in HBM)",
        "eager_ms_per_step": round(e32_t * 1e3, 4),
        "host_buffer_apply_ms": round(t_host32 * 1e3, 3),
        "fp32_rel_l2_err": e32,
        "fp32_over_fp64_speedup": round(t64 / t32, 3),
        "build_s": round(build_s, 2),
        "build_split_s": {"tree": round(t_tree, 2), "skel": round(t_skel, 2), "lists_upload": round(t_setup, 2)},
        "gpu_build": gpu_build,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=None, help="point count of the selected workload")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--leaf", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-fraction", type=float, default=0.02)
    ap.add_argument("--ncheck", type=int, default=1000)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 512 MiB device-to-device copy first (calibrates rocprofv3 byte counters)")
    ap.add_argument("--workload", choices=("laplace3d_cube", "annuli", "sphere", "mixture"), default="laplace3d_cube",
                    help="laplace3d_cube: the headline line (default, with the other workloads as extra keys); "
                         "annuli / sphere / mixture: one workload alone")
    ap.add_argument("--only-main", action="store_true", help="headline workload only (no extra workloads)")
    ap.add_argument("--double-layer", action="store_true", help="sphere workload: Helmholtz double layer")
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--no-tune", action="store_true", help="skip the near-field autotuner (on by default)")
    ap.add_argument("--max-sources", type=int, default=0, help="headline workload: pin the source-run length")
    ap.add_argument("--sym-unroll", type=int, default=0, help="headline workload: pin the symmetric unroll")
    ap.add_argument("--self-block", type=int, default=0,
                    help="headline workload: pin the block of the self segments' symmetric evaluation")
    ap.add_argument("--no-gpu-build", action="store_true",
                    help="laplace3d_cube: skip the second build with the IDs on the GPU")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "annuli":
        out = run_annuli(args, args.steps, args.warmup, not args.no_cpu)
        print(json.dumps(out), flush=True)
        return
    if args.workload in ("sphere", "mixture"):
        out = run_helmholtz(args, args.workload, args.double_layer, args.steps, args.warmup, not args.no_cpu,
                            n=args.n)
        print(json.dumps(out), flush=True)
        return
    if args.n is None:
        args.n = 1_000_000
    line = run_cube(args)
    if not args.only_main:
        import torch

        more = {}
        jobs = [("annuli_32x64rhs", lambda: run_annuli(args, args.steps, args.warmup, False)),
                ("mixture_2M", lambda: run_helmholtz(args, "mixture", False, args.steps, args.warmup, False)),
                ("sphere_1M", lambda: run_helmholtz(args, "sphere", False, args.steps, args.warmup, False)),
                ("sphere_1M_double_layer", lambda: run_helmholtz(args, "sphere", True, args.steps, args.warmup, False))]
        for key, job in jobs:
            ts = time.perf_counter()
            try:
                more[key] = job()
            except Exception as exc:  # one workload's failure must not hide the headline line
                more[key] = {"error": f"{type(exc).__name__}: {exc}"}
                log(f"{key} failed: {exc}")
            more[key]["wall_s"] = round(time.perf_counter() - ts, 1)
            torch.cuda.empty_cache()
        line["more_workloads"] = more
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
