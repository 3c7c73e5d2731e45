"""Launch-parameter autotuner for the device apply.

For one uploaded build and one vector width it times, with HIP events, the
near-field launch over

* the source-run length of a work item (``max_sources``: how many source points
  one 64-target work item covers before the target block is split across
  workgroups -- trades load balance against atomics and re-staging), and
* for FP32 real kernels with at least 16 channels, the matrix-core kernel
  against the VALU kernel (RHS blocking: 64 channels per MFMA block, or the
  widest VALU channel block that fits LDS), and
* the launch order of the work items (heaviest first, or spatial runs per XCD so that
  neighbouring items share an L2),
* where the symmetric near-field kernel runs (one channel, single-layer kernel),
  the number of sources per step of its inner loop (1, 2 or 4: instruction-level
  parallelism against registers, i.e. waves per SIMD),
* with the symmetric kernel in use, the block size of the self segments' symmetric evaluation
  (``SELF_BLOCKS``, or none: self segments one-sided), timed with the best setting of the above,
* per tree level and direction, the launch shape of the interpolation stage
  (``tune_levels``: vector loads in flight per lane for one channel, matrix cores or
  VALU GEMM for >= 16 real FP32 channels), each stage timed alone,

keeps the fastest and caches the choice per (device name, operator, precision,
width, slots, interactions) in a JSON file (``$SKELFMM_AMD_TUNE_CACHE``, default
``~/.cache/skelfmm_amd/autotune.json``). ``tune`` leaves the winning setting
applied to the object.
"""

from __future__ import annotations

import json
import os

import torch

from .layout import ITEM_ORDERS

SPLITS = (1024, 2048, 4096, 8192, 16384, 32768, 65536)
# block sizes of the self segments' symmetric evaluation tried (0, one-sided, is always tried)
SELF_BLOCKS = (128, 192, 256)


def cache_path():
    return os.environ.get("SKELFMM_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "skelfmm_amd", "autotune.json"))


def signature(dev, width):
    return "|".join(str(x) for x in (torch.cuda.get_device_name(dev.device), dev.operator.name, dev.precision,
                                     width, dev.nslots, dev.interactions))


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, db):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def apply_setting(dev, setting):
    dev.set_mfma_min_channels(int(setting.get("mfma_min_channels", 0)))
    if hasattr(dev, "set_sym_unroll"):
        dev.set_sym_unroll(int(setting.get("sym_unroll", 0)))
    if hasattr(dev, "set_self_block"):
        dev.set_self_block(int(setting.get("self_block", 0)))
    if int(setting["max_sources"]) != int(getattr(dev.layout, "max_sources", -1)):
        dev.set_near_field_split(int(setting["max_sources"]))
    order = setting.get("item_order", "heavy")
    if hasattr(dev, "set_item_order") and order != getattr(dev, "item_order", "heavy"):
        dev.set_item_order(order)


def level_variant_choices(dev, width):
    """Launch variants of the interpolation stages at ``width`` channels: loads in flight for one
    channel (3 shapes), matrix cores or VALU GEMM for >= 16 real FP32 channels, else one."""
    nch = width // 2 if dev.complex_vectors else width
    if nch == 1:
        return (0, 1, 2)
    if nch >= 16 and dev.precision == 1 and not dev.complex_t:
        return (0, 1)
    return (0,)


def tune_levels(dev, width=1, reps=3):
    """Per upward and downward stage, the fastest launch variant at ``width`` channels (each
    stage timed alone with HIP events); applies it and returns the choice with the timings."""
    choices = level_variant_choices(dev, width)
    timed = {}
    for v in choices:
        dev.set_level_variants(width, up=[v] * len(dev.layout.up), down=[v] * len(dev.layout.down))
        timed[v] = dev.time_levels(width, reps=reps)
    best, default_ms, tuned_ms = {}, 0.0, 0.0
    for name in ("up", "down"):
        n = len(timed[choices[0]][name])
        pick = [min(choices, key=lambda v: timed[v][name][i]) for i in range(n)]
        best[name] = pick
        default_ms += sum(timed[0][name])
        tuned_ms += sum(timed[p][name][i] for i, p in enumerate(pick))
    dev.set_level_variants(width, up=best["up"], down=best["down"])
    return {"up": best["up"], "down": best["down"], "levels_default_ms": round(default_ms, 4),
            "levels_tuned_ms": round(tuned_ms, 4),
            "levels_ms": {str(v): {k: [round(t, 4) for t in ts] for k, ts in timed[v].items()} for v in choices}}


def tune(dev, width=1, reps=3, use_cache=True, splits=SPLITS, levels=True):
    """Pick and apply the fastest near-field setting for ``dev`` at ``width`` channels, and (with
    ``levels``) the launch variant of every interpolation stage; returns the setting."""
    path = cache_path()
    key = signature(dev, width)
    db = _load(path) if use_cache else {}
    if key in db:
        apply_setting(dev, db[key])
        if "levels" in db[key]:
            dev.set_level_variants(width, up=db[key]["levels"]["up"], down=db[key]["levels"]["down"])
        return dict(db[key], cached=True)
    nch = width // 2 if dev.complex_vectors else width
    mfma_opts = [0]
    if dev.precision == 1 and not dev.operator.is_complex and nch >= 16:
        mfma_opts = [0, -1]
    sym = getattr(dev, "use_symmetric", False) and nch == 1
    unrolls = (1, 2, 4) if sym else (0,)
    orders = ITEM_ORDERS if hasattr(dev, "set_item_order") else ("heavy",)
    results = []
    for order in orders:
        if hasattr(dev, "set_item_order"):
            dev.set_item_order(order)
        for mf in mfma_opts:
            dev.set_mfma_min_channels(mf)
            for un in unrolls:
                if hasattr(dev, "set_sym_unroll"):
                    dev.set_sym_unroll(un)
                for ms in splits:
                    dev.set_near_field_split(ms)
                    t = dev.time_near_field(width, reps=reps)
                    results.append((t, ms, mf, un, order))
    t, ms, mf, un, order = min(results)
    best = {"max_sources": ms, "mfma_min_channels": mf, "sym_unroll": un, "item_order": order,
            "near_field_ms": round(t, 4), "tried": [[round(a, 4), b, c, d, e] for a, b, c, d, e in results]}
    apply_setting(dev, best)
    if sym and hasattr(dev, "set_self_block"):
        # the candidates differ by a few percent: each is timed twice, the faster run kept
        timed = []
        for blk in (0,) + tuple(SELF_BLOCKS):
            dev.set_self_block(blk)
            timed.append((min(dev.time_near_field(width, reps=max(reps, 5)) for _ in range(2)), blk))
        t, blk = min(timed)
        best["self_block"] = blk
        best["near_field_ms"] = round(t, 4)
        best["self_block_tried"] = [[round(a, 4), b] for a, b in timed]
        apply_setting(dev, best)
    if levels and (len(dev.layout.up) or len(dev.layout.down)):
        best["levels"] = tune_levels(dev, width, reps=reps)
    if use_cache:
        db[key] = best
        _store(path, db)
    return dict(best, cached=False)
