"""Serialized built operator: one versioned ``.npz`` file with everything the device apply needs.

The reference keeps a build only in memory (``skelfmm/skel_fmm.py:182-186``). This format
stores the flattened box layout of a build (``layout.BoxLayout``: slot tables, points,
interpolation matrices, the upward and downward level stages, the near-field lists and
their symmetric split) together with the operator, its parameter, the tolerance and the
points in the caller's order, so a build can be reused across runs and processes without
the tree, the interpolative decompositions or the lists being formed again.

File layout (``numpy.savez``, loaded with ``allow_pickle=False``, so opening a file
executes nothing from it):

* ``__header__``: UTF-8 JSON as a ``uint8`` array: ``format`` (``FORMAT``), ``version``
  (``VERSION``), operator name, ``kappa``, ``tol``, ``leaf_size``, sizes, the scalar fields
  of the layout and of every level stage and near-field list, and a CRC-32 per array;
* ``layout/<field>``, ``up/<i>/<field>``, ``down/<i>/<field>``, ``split/<i>/<field>``:
  the layout's arrays;
* ``points`` (``(N, d)``, caller's order) and, for geometry-aware operators, ``geometry``;
* ``extra/<name>``: per-box build tables the caller adds (``rank_list``, ``bs_list``).

A file of another format name or a newer version is rejected with ``ValueError``; so is
an array whose CRC-32 does not match (``verify=True``), and so is a layout whose index tables
leave the arrays they index (``layout.check_layout``, always run: the CRC lives in the same
file, so it only catches accidental corruption).
"""

from __future__ import annotations

import json
import zlib
from typing import NamedTuple

import numpy as np

from .layout import BoxLayout, LevelStage, NearLists, check_layout

__all__ = ["FORMAT", "VERSION", "Serialized", "save_layout", "load_layout"]

FORMAT = "skelfmm_amd.built_operator"
VERSION = 1

_STAGE_ARRAYS = ("boxes", "box_off", "box_bs", "box_rank", "box_toff", "gather_dst", "gather_src", "items")
_LIST_ARRAYS = NearLists.FIELDS + ("items", "item_cost")
_LAYOUT_ARRAYS = ("off", "slot_point", "slot_box", "points", "normals", "centers", "leaf_slot", "leaf_user", "T",
                  "toff") + _LIST_ARRAYS
_LAYOUT_SCALARS = ("ndim", "N", "nboxes", "root_level", "nslots", "interactions", "evaluations", "max_sources")


class Serialized(NamedTuple):
    header: dict
    layout: BoxLayout
    points: np.ndarray
    geometry: np.ndarray | None
    extra: dict


def _crc(a):
    return zlib.crc32(np.ascontiguousarray(a).view(np.uint8).reshape(-1)) & 0xFFFFFFFF


def _scalar(v):
    if isinstance(v, (np.integer,)):
        return int(v)
    if isinstance(v, (np.floating,)):
        return float(v)
    return v


def save_layout(path, L, *, operator_name, kappa, tol, leaf_size, points, geometry=None, extra=None,
                extra_arrays=None):
    """Write layout ``L`` and its operator metadata to ``path`` (a ``.npz`` file or file object)."""
    arrays = {}
    header = {"format": FORMAT, "version": VERSION, "operator": str(operator_name), "kappa": float(kappa),
              "tol": float(tol), "leaf_size": int(leaf_size), "layout": {}, "up": [], "down": [], "split": None,
              "extra": extra or {}}
    for k in _LAYOUT_SCALARS:
        if hasattr(L, k):
            header["layout"][k] = _scalar(getattr(L, k))
    for k in _LAYOUT_ARRAYS:
        v = getattr(L, k, None)
        if v is not None:
            arrays["layout/" + k] = v.cpu().numpy() if hasattr(v, "cpu") else np.asarray(v)
    for name in ("up", "down"):
        for i, S in enumerate(getattr(L, name)):
            header[name].append({"level": int(S.level)})
            for k in _STAGE_ARRAYS:
                arrays[f"{name}/{i}/{k}"] = np.asarray(getattr(S, k))
    if getattr(L, "split", None) is not None:
        header["split"] = []
        for i, part in enumerate(L.split):
            header["split"].append({"tpi": int(part.tpi), "interactions": int(part.interactions)})
            for k in _LIST_ARRAYS:
                arrays[f"split/{i}/{k}"] = np.asarray(getattr(part, k))
    arrays["points"] = np.ascontiguousarray(points)
    if geometry is not None:
        arrays["geometry"] = np.ascontiguousarray(geometry)
    for k, v in (extra_arrays or {}).items():
        arrays["extra/" + k] = np.ascontiguousarray(v)
    header["crc32"] = {k: _crc(v) for k, v in arrays.items()}
    arrays["__header__"] = np.frombuffer(json.dumps(header).encode("utf-8"), dtype=np.uint8)
    if hasattr(path, "write"):
        np.savez(path, **arrays)
    else:
        with open(path, "wb") as fh:
            np.savez(fh, **arrays)


def load_layout(path, verify=True):
    """Read a file written by ``save_layout`` into a ``Serialized`` record."""
    with np.load(path, allow_pickle=False) as z:
        if "__header__" not in z.files:
            raise ValueError(f"{path}: not a serialized built operator (no header)")
        header = json.loads(bytes(z["__header__"]).decode("utf-8"))
        if header.get("format") != FORMAT:
            raise ValueError(f"{path}: format {header.get('format')!r}, expected {FORMAT!r}")
        if int(header.get("version", -1)) > VERSION or int(header.get("version", -1)) < 1:
            raise ValueError(f"{path}: format version {header.get('version')} not supported (this library reads "
                             f"versions 1..{VERSION})")
        arrays = {k: z[k] for k in z.files if k != "__header__"}
    if verify:
        crcs = header.get("crc32", {})
        if set(crcs) != set(arrays):
            raise ValueError(f"{path}: array set does not match the header")
        for k, v in arrays.items():
            if _crc(v) != crcs[k]:
                raise ValueError(f"{path}: checksum mismatch in {k!r}")
    L = BoxLayout()
    for k, v in header["layout"].items():
        setattr(L, k, v)
    for k in _LAYOUT_ARRAYS:
        setattr(L, k, arrays.get("layout/" + k))
    for name in ("up", "down"):
        stages = []
        for i, meta in enumerate(header[name]):
            a = [arrays[f"{name}/{i}/{k}"] for k in _STAGE_ARRAYS]
            boxes, rest = a[0], a[1:]
            stages.append(LevelStage(int(meta["level"]), boxes, *rest))
        setattr(L, name, stages)
    L.split = None
    if header["split"] is not None:
        parts = []
        for i, meta in enumerate(header["split"]):
            part = NearLists()
            for k in _LIST_ARRAYS:
                setattr(part, k, arrays[f"split/{i}/{k}"])
            part.tpi = int(meta["tpi"])
            part.interactions = int(meta["interactions"])
            parts.append(part)
        L.split = tuple(parts)
    extra = {k[len("extra/"):]: v for k, v in arrays.items() if k.startswith("extra/")}
    try:
        check_layout(L)
    except (AttributeError, TypeError, IndexError) as exc:  # a missing or malformed table
        raise ValueError(f"{path}: incomplete box layout ({exc})") from exc
    except ValueError as exc:
        raise ValueError(f"{path}: {exc}") from exc
    points = arrays["points"]
    if points.shape != (L.N, L.ndim):
        raise ValueError(f"{path}: points shape {points.shape}, expected {(L.N, L.ndim)}")
    return Serialized(header, L, points, arrays.get("geometry"), extra)
