"""Apply throughput of the skeletonized FMM on one MI355X.

Workload (default): 3D Laplace (LAPLACE_3D), N = 1,000,000 points i.i.d. uniform
on [0,1]^3 (numpy default_rng(0)), charges RandomState(1).randn, tol 1e-6,
leaf_size 400, one right-hand side. A step is one apply (charges resident in
HBM, result left in HBM). The reported value is the FP32 precision form in
Mpoints*RHS/s; the plain FP64 form, the FP32/FP64 speed-up, relative l2 errors
of both forms against an FP64 direct sum at 1,000 sampled targets and the build
seconds are reported beside it.

``roofline`` is for the near-field kernel (the dominant one): algorithmic
flops per launch (interactions x flops per interaction) over its mean launch
duration, timed with HIP events on the stream it runs on.
``cpu_baseline`` is the padded torch-CPU restatement of the reference apply run on
two sampled fractions of every interaction list (with full interpolation passes) and
extrapolated by a linear fit in the interaction count.

Other workloads: ``--workload annuli``: 32 independent curvy annuli (2D Laplace,
N_i log-uniform in [1e4, 5e5], tol 1e-5, leaf 128) with 64 float32 right-hand
sides each, applied together as one graph-captured apply; ``sphere``: 1M points on
the unit sphere, 3D Helmholtz kappa 10 pi, tol 1e-5, leaf 256 (``--double-layer``
for the double layer), GPU build; ``mixture``: 2M points of a 16-component Gaussian
mixture, 3D Helmholtz kappa 20, tol 1e-6, leaf 256, built twice on the GPU
(adaptive IDs, then learned-rank sketched IDs), complex64 apply.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu] [--no-tune]
       python bench.py --workload annuli|sphere|mixture
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

# flops per interaction (FMA = 2): dx,dy,dz (3) + r^2 (mul + 2 fma = 5) + rsqrt (1) + q*K accumulate (2).
# The symmetric near field evaluates K once for both directions of a pair, so flops per launch are
# counted as evaluations x (flops - ACC_FLOPS) + interactions x ACC_FLOPS.
ACC_FLOPS = 2
FLOPS_PER_INTERACTION = {"laplace_3d": 11, "laplace_2d": 11,
                         # + r = r^2 rsqrt (1), kappa r (1), sin and cos (2), e^{ikr}/r (2), complex q K (8)
                         "helmholtz_3d": 23,
                         # + (x - y).n (5) and the (ikr - 1) r^-2 factor (4)
                         "helmholtz_double_layer_3d": 32}
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
# HBM bytes per launch of the dominant FP32 near-field kernel on the default workload (one-sided
# kernel, or the symmetric kernel: its FP64 reverse-sum atomics count as writes), from the two
# rocprofv3 --pmc passes summarised in profiles/apply_laplace3d_uniform1M_hbm_traffic.md
# (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, scale factors from a 512 MiB copy of the same run)
P2P_FP32_TRAFFIC_BYTES = {(1_000_000, 1e-6, 400, False): 2549.1e6 + 14.2e6,
                          (1_000_000, 1e-6, 400, True): 641.3e6 + 631.0e6}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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


def run_annuli(args):
    """Batched uneven 2D Laplace: every instance and RHS in one graph-captured apply."""
    import torch

    from oracle.direct import direct_sum_c
    from skelfmm_amd import autotune, hip, operators
    from skelfmm_amd.batched import BatchedInstances
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.tree import BalancedTree

    hip.require_gpu()
    dev = torch.device("cuda", 0)
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
    log(f"annuli: {len(ns)} instances, {ntot} points, {args.nrhs} RHS; build {build_s:.1f}s (host {t_host:.1f}); "
        f"interactions/apply {d32.interactions:.3e}")
    width = args.nrhs
    tuned = None
    if not args.no_tune:
        tuned = autotune.tune(d32, width)
        log(f"autotune FP32: {tuned}")
    x32 = d32.concat([torch.as_tensor(q) for q in qs])
    x64 = x32.to(torch.float64)

    def graph_time(d, x):
        out = torch.empty_like(x)
        g = d.capture(x, out)
        for _ in range(args.warmup):
            g.replay()
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        return (time.perf_counter() - ts) / args.steps, out

    t32, y32 = graph_time(d32, x32)
    t64, y64 = graph_time(d64, x64)
    p32 = d32.time_near_field(width, reps=5) * 1e-3
    log(f"apply (graph replay) FP32 {t32 * 1e3:.2f} ms, FP64 {t64 * 1e3:.2f} ms; near field FP32 {p32 * 1e3:.2f} ms")

    # per-instance relative l2 error on 4 RHS columns at 100 sampled targets vs the FP64 direct sum
    errs32, errs64 = [], []
    parts32 = d32.split(y32)
    parts64 = d64.split(y64)
    for i, (P, q) in enumerate(zip(pts, qs)):
        idx = np.random.default_rng(i).choice(P.shape[0], min(100, P.shape[0]), replace=False)
        ref = direct_sum_c(op.name, P[idx], P, q[:, :4].astype(np.float64))
        a = parts32[i][idx, :4].cpu().numpy().astype(np.float64)
        b = parts64[i][idx, :4].cpu().numpy()
        errs32.append(float(np.linalg.norm(a - ref) / np.linalg.norm(ref)))
        errs64.append(float(np.linalg.norm(b - ref) / np.linalg.norm(ref)))
    log(f"rel-l2 vs direct sum (max over instances): FP32 {max(errs32):.3e}, FP64 {max(errs64):.3e}")

    mfma_tf = 2.0 * d32.interactions * width / p32 / 1e12
    cpu = None
    if not args.no_cpu:
        from oracle.batched_apply import BatchedApplyCPU

        ncores = min(16, len(os.sched_getaffinity(0)))
        torch.set_num_threads(ncores)
        k = int(np.argmin(ns))
        ba = BatchedApplyCPU(builds[k], bytes_budget=10e9, pair_fraction=0.25)
        q0 = qs[k].astype(np.float64)
        ba.matvec(q0)
        tc = time.perf_counter()
        ba.matvec(q0)
        el = time.perf_counter() - tc
        est = el * d64.interactions / max(ba.interactions, 1)
        cpu = {"value": round(ntot * width / est / 1e6, 6), "unit": "Mpoints*RHS/s", "cores": ncores,
               "kind": "port",
               "sample": (f"padded torch-CPU FP64 restatement of the reference apply, {width} RHS, on 25% of the pairs "
                          f"of the smallest instance (N={int(ns[k])}, {ba.interactions:.3e} interactions, {el:.1f}s), "
                          f"extrapolated to {d64.interactions:.3e} interactions")}
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "Mpoints*RHS/s per apply, FP32 precision form (FP64 form beside it)",
        "value": round(ntot * width / t32 / 1e6, 3),
        "unit": "Mpoints*RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 32 curvy annuli, N_i log-uniform in [1e4, 5e5] (default_rng(2)), float32 N(0,1) charges",
        "config": {"workload": "batched_annuli_laplace_2d", "instances": int(len(ns)), "N_total": ntot,
                   "nrhs": width, "tol": tol, "leaf_size": leaf, "graph": True,
                   "interactions_per_apply": int(d32.interactions), "autotune": tuned},
        "fp64": {"value": round(ntot * width / t64 / 1e6, 3), "ms_per_step": round(t64 * 1e3, 4),
                 "max_rel_l2_err": max(errs64)},
        "fp32_max_rel_l2_err": max(errs32),
        "build_s": round(build_s, 2),
        "roofline": {"bound": "mfma", "achieved": round(mfma_tf, 3), "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                     "frac": round(mfma_tf / PEAK_TFLOPS["fp32"], 4), "traffic": None,
                     "kernel": "near field (matrix-core kernel at >= 16 RHS unless the autotuner picked VALU)",
                     "kernel_ms": round(p32 * 1e3, 4), "flops_counted": "2 x interactions x RHS (K q products)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


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


def run_helmholtz(args):
    """3D Helmholtz workloads: ``sphere`` (1M points on the unit sphere, kappa 10 pi, tol 1e-5,
    leaf 256, single or double layer) and ``mixture`` (2M-point Gaussian mixture, kappa 20,
    tol 1e-6, leaf 256) whose build runs twice on the GPU: adaptive IDs, then learned-rank
    sketched IDs with adaptive fallback. Value: complex64 apply throughput."""
    import torch

    from oracle.direct import direct_sum_c
    from oracle.kernels import GEOM_SIDE
    from skelfmm_amd import hip, learned_rank, operators
    from skelfmm_amd.batched import SkelFMMBatched
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.tree import BalancedTree

    hip.require_gpu()
    dev = torch.device("cuda", 0)
    extra = {}
    if args.workload == "sphere":
        N = args.n
        v = np.random.default_rng(0).normal(size=(N, 3))
        X = v / np.linalg.norm(v, axis=1, keepdims=True)
        op = operators.HELMHOLTZ_DOUBLE_LAYER_3D if args.double_layer else operators.HELMHOLTZ_3D
        geom = X.copy() if args.double_layer else None
        kappa, tol, leaf = 10 * np.pi, 1e-5, 256
        workload = "helmholtz_3d_sphere" + ("_double_layer" if args.double_layer else "")
        data = "synthetic: unit-sphere points default_rng(0) normal/|.|, normals = points"
    else:
        N = 2_000_000 if args.n == 1_000_000 else args.n
        X, clipped = mixture_points(N)
        op, geom = operators.HELMHOLTZ_3D, None
        kappa, tol, leaf = 20.0, 1e-6, 256
        workload = "helmholtz_3d_gaussian_mixture"
        data = f"synthetic: 16-component Gaussian mixture default_rng(3), sigma 0.02, {clipped} points clipped to [0,1]^3"
        extra["clipped_points"] = clipped
    rs = np.random.RandomState(1)
    q = rs.randn(N) + 1j * rs.randn(N)

    t0 = time.perf_counter()
    tree = BalancedTree(X, leaf_size=leaf)
    t_tree = time.perf_counter() - t0
    log(f"{workload}: N={N}, tree {t_tree:.1f}s, levels {tree.nlevels}, boxes {tree.nboxes}")

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

    builds = {}
    f_a, d_a, tid_a, tset_a = build(None)
    builds["adaptive"] = {"id_s": round(tid_a, 3), "lists_upload_s": round(tset_a, 2),
                          "build_s": round(t_tree + tid_a + tset_a, 2), "rank_sum": int(f_a.rank_list.sum()),
                          "level_s": {int(k): round(v, 3) for k, v in f_a.level_times.items()}}
    log(f"adaptive GPU build: IDs {tid_a:.2f}s, lists+upload {tset_a:.1f}s, root level {f_a.root_level}")
    f_use, d64 = f_a, d_a
    model = learned_rank.load_default() if args.workload == "mixture" else None
    if model is not None:
        # warm the sketch path's library calls once outside the timing (rocBLAS handle set-up)
        f_l, d_l, tid_l, tset_l = build(model)
        st = f_l.gpu_stats
        builds["learned"] = {"id_s": round(tid_l, 3), "lists_upload_s": round(tset_l, 2),
                             "build_s": round(t_tree + tid_l + tset_l, 2), "rank_sum": int(f_l.rank_list.sum()),
                             "stats": st, "fallback_rate": round(st["fallback"] / max(1, st["sketched"] + st["fallback"]), 4),
                             "rejected_by": dict(f_l.gpu_skel.sketch_stats),
                             "level_s": {int(k): round(v, 3) for k, v in f_l.level_times.items()}}
        builds["id_speedup_learned_over_adaptive"] = round(tid_a / tid_l, 3)
        log(f"learned GPU build: IDs {tid_l:.2f}s ({tid_a / tid_l:.2f}x), stats {st}")
        f_use, d64 = f_l, d_l
    d32 = SkelFMMBatched(tree, op, kappa=kappa).build_from_copy(f_use, use_single_precision=True)

    q128 = torch.as_tensor(q, dtype=torch.complex128, device=dev)
    q64 = q128.to(torch.complex64)

    def timed(d, x):
        y = torch.empty_like(x)
        for _ in range(args.warmup):
            d.apply_device(x, out=y)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(args.steps):
            d.apply_device(x, out=y)
        torch.cuda.synchronize()
        return (time.perf_counter() - ts) / args.steps, y

    t64, y64 = timed(d32, q64)
    t128, y128 = timed(d64, q128)
    p64 = d32.time_near_field(2, reps=5) * 1e-3
    p128 = d64.time_near_field(2, reps=5) * 1e-3
    log(f"apply complex64 {t64 * 1e3:.2f} ms, complex128 {t128 * 1e3:.2f} ms; near field {p64 * 1e3:.2f} / "
        f"{p128 * 1e3:.2f} ms")

    idx = np.random.default_rng(0).choice(N, min(args.ncheck, N), replace=False)
    side = GEOM_SIDE.get(op.name)
    ref = direct_sum_c(op.name, X[idx], X, q, kappa=kappa, geom_trg=geom[idx] if side == "target" else None,
                       geom_src=geom if side == "source" else None)
    err = lambda y: float(np.linalg.norm(y.cpu().numpy()[idx] - ref) / np.linalg.norm(ref))
    e64, e128 = err(y64), err(y128)
    errs = {"complex64": e64, "complex128": e128}
    if "learned" in builds:
        _, ya = timed(d_a, q128)
        errs["complex128_adaptive_build"] = err(ya)
    log(f"rel-l2 vs direct sum at {idx.size} targets: {errs}")

    fpi = FLOPS_PER_INTERACTION[op.name]
    tf = d32.interactions * fpi / p64 / 1e12
    line = {
        "metric": "Mpoints*RHS/s per apply, complex64 precision form (complex128 form and build s beside it)",
        "value": round(N / t64 / 1e6, 3),
        "unit": "Mpoints*RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t64 * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": data + "; charges RandomState(1) randn + 1j randn",
        "config": {"workload": workload, "N": N, "kappa": kappa, "tol": tol, "leaf_size": leaf, "nrhs": 1,
                   "nlevels": int(tree.nlevels), "root_level": int(f_use.root_level),
                   "interactions_per_apply": int(d32.interactions), **extra},
        "complex128": {"value": round(N / t128 / 1e6, 3), "ms_per_step": round(t128 * 1e3, 4)},
        "rel_l2_err": errs,
        "builds": builds,
        "roofline": {"bound": "mfma", "achieved": round(tf, 3), "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                     "frac": round(tf / PEAK_TFLOPS["fp32"], 4), "traffic": None,
                     "compute_unit": "VALU (per-pair kernel evaluation with sin/cos)",
                     "kernel": f"near field complex64 ({op.name})", "kernel_ms": round(p64 * 1e3, 4),
                     "flops_per_interaction": fpi},
        "cpu_baseline": _cpu_baseline(f_use, q, d64.interactions, N, args) if not args.no_cpu else None,
    }
    print(json.dumps(line), flush=True)


def cpu_apply_estimate(fmm, q, interactions, fraction):
    """Seconds of one full CPU apply, from two sampled runs of the oracle's padded apply.

    Runs on ``fraction`` and ``2 fraction`` of the pairs of every list (interpolation passes in
    full both times) and fits t = a + b I over the sampled interaction counts I; the estimate is
    a + b x ``interactions``. Returns (seconds, sample description)."""
    from oracle.batched_apply import BatchedApplyCPU

    pts = []
    for f in (fraction, 2 * fraction):
        ba = BatchedApplyCPU(fmm, bytes_budget=10e9, pair_fraction=f)
        if not pts:
            ba.matvec(q)  # warm-up
        tc = time.perf_counter()
        ba.matvec(q)
        pts.append((ba.interactions, time.perf_counter() - tc))
    (i1, t1), (i2, t2) = pts
    b = max(t2 - t1, 0.0) / max(i2 - i1, 1)
    a = max(t1 - b * i1, 0.0)
    est = a + b * interactions
    desc = (f"padded torch-CPU restatement of the reference apply on {fraction:.0%} and {2 * fraction:.0%} of the "
            f"pairs of every interaction list ({i1:.3e} / {i2:.3e} interactions, {t1:.1f} / {t2:.1f} s) with full "
            f"interpolation passes; linear fit t = a + b I (a = {a:.2f} s), evaluated at {interactions:.3e} "
            f"interactions")
    return est, desc


def _cpu_baseline(fmm, q, interactions, N, args, fraction=0.01):
    """Padded torch-CPU restatement of the reference apply on a sample of every pair list."""
    import torch

    ncores = min(16, len(os.sched_getaffinity(0)))
    torch.set_num_threads(ncores)
    est, desc = cpu_apply_estimate(fmm, q, interactions, fraction)
    cpu = {"value": round(N / est / 1e6, 6), "unit": "Mpoints*RHS/s", "cores": ncores, "kind": "port",
           "sample": "complex128 " + desc}
    log(f"cpu baseline: {cpu}")
    return cpu


def _p2p_roofline(d32, d64, s32, s64, st32, st64, p32, p64, tf32, tf64, fpi, N, args):
    """Roofline of the dominant kernel: the symmetric near field when it runs, else the one-sided one."""
    sym = bool(d32.use_symmetric and s32)
    ach, ms = (st32, s32) if sym else (tf32, p32)
    out = {
        "bound": "mfma",
        "achieved": round(ach, 3),
        "peak": PEAK_TFLOPS["fp32"],
        "unit": "TFLOP/s",
        "frac": round(ach / PEAK_TFLOPS["fp32"], 4),
        "traffic": P2P_FP32_TRAFFIC_BYTES.get((N, args.tol, args.leaf, sym)),
        "traffic_unit": "bytes per launch (HBM read + write) of the dominant kernel",
        "compute_unit": "VALU (per-pair kernel evaluation; no matrix cores at 1 RHS)",
        "kernel": ("p2p_sym_kernel<float, LAPLACE_3D> (pairs of lists, both directions; zero fill and fold "
                   "of the reverse sums included)" if sym else "p2p_kernel<float, LAPLACE_3D> (near field)"),
        "kernel_ms": round(ms * 1e3, 4),
        "flops_counted": (f"pairs x {fpi - ACC_FLOPS} + 2 pairs x {ACC_FLOPS} ({d32.sym_pairs} pairs)" if sym else
                          f"interactions x {fpi}"),
        "near_field_call": {"achieved": round(tf32, 3), "ms": round(p32 * 1e3, 4),
                            "evaluations": int(d32.evaluations if d32.use_symmetric else d32.interactions),
                            "interactions": int(d32.interactions)},
        "fp64_kernel": {"achieved": round(st64 if (sym and st64) else tf64, 3), "peak": PEAK_TFLOPS["fp64"],
                        "frac": round((st64 if (sym and st64) else tf64) / PEAK_TFLOPS["fp64"], 4),
                        "kernel_ms": round((s64 if (sym and s64) else p64) * 1e3, 4),
                        "near_field_call_ms": round(p64 * 1e3, 4)},
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--leaf", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-fraction", type=float, default=0.02)
    ap.add_argument("--ncheck", type=int, default=1000)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 512 MiB device-to-device copy first (calibrates rocprofv3 byte counters)")
    ap.add_argument("--workload", choices=("laplace3d_cube", "annuli", "sphere", "mixture"), default="laplace3d_cube",
                    help="laplace3d_cube: 1M-point 3D Laplace (default); annuli: 32 batched 2D annuli x 64 RHS; "
                         "sphere: 1M-point 3D Helmholtz on the sphere; mixture: 2M-point Gaussian mixture, "
                         "learned-rank vs adaptive GPU build")
    ap.add_argument("--double-layer", action="store_true", help="sphere workload: Helmholtz double layer")
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--no-tune", action="store_true", help="skip the near-field autotuner (on by default)")
    ap.add_argument("--no-gpu-build", action="store_true",
                    help="laplace3d_cube: skip the second build with the IDs on the GPU")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "annuli":
        return run_annuli(args)
    if args.workload in ("sphere", "mixture"):
        return run_helmholtz(args)

    import torch

    from skelfmm_amd import hip, operators
    from skelfmm_amd.batched import SkelFMMBatched
    from skelfmm_amd.build import RecursiveSkeletonization
    from skelfmm_amd.tree import BalancedTree

    if not os.path.exists(hip.LIB_PATH):
        import __graft_entry__

        __graft_entry__.build()
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
    if not args.no_tune:
        from skelfmm_amd import autotune

        tuned = {"fp64": autotune.tune(d64, 1), "fp32": autotune.tune(d32, 1)}
        log(f"autotune: {tuned}")
    q64 = torch.as_tensor(q, dtype=torch.float64, device=dev)
    q32 = q64.to(torch.float32)

    def timed(devapply, x):
        """(graph-replay s per apply, eager s per apply, result): K steps each after W warmups."""
        y = torch.empty_like(x)
        for _ in range(args.warmup):
            devapply.apply_device(x, out=y)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(args.steps):
            devapply.apply_device(x, out=y)
        torch.cuda.synchronize()
        t_eager = (time.perf_counter() - ts) / args.steps
        g = devapply.capture(x, y)
        for _ in range(args.warmup):
            g.replay()
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        return (time.perf_counter() - ts) / args.steps, t_eager, y

    t64, e64_t, y64 = timed(d64, q64)
    t32, e32_t, y32 = timed(d32, q32)
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

    # near-field launch alone (the plan's own arguments), HIP events on the stream it runs on
    def p2p_time(devapply, width=1, reps=10):
        return devapply.time_near_field(width, reps=reps) * 1e-3

    p64 = p2p_time(d64)
    p32 = p2p_time(d32)
    # dominant kernel alone: the symmetric near field (zero fill + p2p_sym_kernel + fold)
    s32 = d32.time_near_field(1, reps=10, part="sym") * 1e-3 if d32.use_symmetric else None
    s64 = d64.time_near_field(1, reps=10, part="sym") * 1e-3 if d64.use_symmetric else None
    fpi = FLOPS_PER_INTERACTION[op.name]

    def nf_flops(d):
        ev = d.evaluations if d.use_symmetric else d.interactions
        return ev * (fpi - ACC_FLOPS) + d.interactions * ACC_FLOPS

    tf64 = nf_flops(d64) / p64 / 1e12
    tf32 = nf_flops(d32) / p32 / 1e12
    # symmetric kernel: pairs x (flops - ACC_FLOPS) + 2 pairs x ACC_FLOPS
    sym_flops = lambda d: d.sym_pairs * (fpi - ACC_FLOPS) + 2 * d.sym_pairs * ACC_FLOPS
    st32 = sym_flops(d32) / s32 / 1e12 if s32 else None
    st64 = sym_flops(d64) / s64 / 1e12 if s64 else None
    log(f"near-field kernel FP64 {p64 * 1e3:.2f} ms ({tf64:.1f} TF/s), FP32 {p32 * 1e3:.2f} ms ({tf32:.1f} TF/s)")

    # errors against the FP64 direct sum at sampled targets
    from oracle.direct import direct_sum_c

    idx = np.random.default_rng(0).choice(N, min(args.ncheck, N), replace=False)
    ref = direct_sum_c(op.name, X[idx], X, q)
    y64h = y64.cpu().numpy()
    y32h = y32.cpu().numpy().astype(np.float64)
    e64 = float(np.linalg.norm(y64h[idx] - ref) / np.linalg.norm(ref))
    e32 = float(np.linalg.norm(y32h[idx] - ref) / np.linalg.norm(ref))
    log(f"rel-l2 vs direct sum at {idx.size} targets: FP64 {e64:.3e}, FP32 {e32:.3e}")

    # the same tree skeletonized with the IDs on the GPU (batched adaptive pivoted QR), timed and
    # checked like the host build
    gpu_build = None
    if not args.no_gpu_build:
        torch.cuda.synchronize()
        tg = time.perf_counter()
        fg = RecursiveSkeletonization(tree, op, kappa=0.0, tol=args.tol)
        fg.skel_tree(npoints_max=500, device="gpu")
        torch.cuda.synchronize()
        t_gid = time.perf_counter() - tg
        tg = time.perf_counter()
        fg.setup_lists()
        dg = SkelFMMBatched(tree, op).build_from_copy(fg, use_single_precision=True)
        torch.cuda.synchronize()
        t_gset = time.perf_counter() - tg
        yg = dg.matvec(q32).cpu().numpy().astype(np.float64)
        eg = float(np.linalg.norm(yg[idx] - ref) / np.linalg.norm(ref))
        gpu_build = {"id_s": round(t_gid, 3), "lists_upload_s": round(t_gset, 2),
                     "build_s": round(t_tree + t_gid + t_gset, 2), "rank_sum": int(fg.rank_list.sum()),
                     "host_rank_sum": int(fmm.rank_list.sum()), "fp32_rel_l2_err": eg}
        log(f"GPU build: IDs {t_gid:.2f}s, lists+upload {t_gset:.2f}s, FP32 apply error {eg:.3e}")
        del dg, fg

    cpu = None
    if not args.no_cpu:
        ncores = min(16, len(os.sched_getaffinity(0)))
        torch.set_num_threads(ncores)
        est, desc = cpu_apply_estimate(fmm, q, d64.interactions, args.cpu_fraction)
        cpu = {
            "value": round(N / est / 1e6, 6),
            "unit": "Mpoints*RHS/s",
            "cores": ncores,
            "kind": "port",
            "sample": "FP64 " + desc,
        }
        log(f"cpu baseline: {cpu}")

    value = N / t32 / 1e6
    line = {
        "metric": "Mpoints*RHS/s per apply, FP32 precision form (FP64 form and build s beside it)",
        "value": round(value, 3),
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
                   "autotune": None if tuned is None else {k: {"max_sources": v["max_sources"]}
                                                           for k, v in tuned.items()}},
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
        "roofline": _p2p_roofline(d32, d64, s32, s64, st32, st64, p32, p64, tf32, tf64, fpi, N, args),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
